"""Build recipe for the CPU oracle (test infrastructure; see qbf_oracle.c).

gcc -O3 -fopenmp on oracle/qbf_oracle.c -> oracle/_build/liboracle.so (git-ignored).
The reference qubo_lib is pure Python (its only native code is CUDA held in Python
strings, compiled by NVRTC at run time), so there is no C/C++ reference source to build
into oracle/_ref; the oracle is pinned by the checks listed in qbf_oracle.c instead.
"""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build" / "liboracle.so"


def build(out: Path = OUT) -> Path:
    out.parent.mkdir(parents=True, exist_ok=True)
    # generic x86-64 code: the same build runs on any x86-64 host
    cmd = ["gcc", "-O3", "-fopenmp", "-fPIC", "-shared", "-o", str(out), str(HERE / "qbf_oracle.c")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("oracle build failed")
    return out


if __name__ == "__main__":
    print(build())
