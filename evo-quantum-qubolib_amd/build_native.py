"""Build step for the native parts of qubo_lib_amd (run by __graft_entry__.build()).

  libqbx.so      hipcc, gfx950: kernels + C ABI (include/qbx.h)  -> qubo_lib_amd/libqbx.so
  libqbx_cpu.so  g++ -fopenmp: the same lane walk on the host CPU, the device='cpu' backend
                 (include/qbx_cpu.h)                              -> qubo_lib_amd/libqbx_cpu.so
  libqbx_emu.so  g++: host build of the same lane walk, loaded only by the CPU test
                 suite to check the walk's indexing without a GPU -> tests/_build/
  qbx_c_check    gcc -std=c99: a plain C caller of the C ABI, run by a GPU test
                 -> tests/_build/

Every call recompiles.  Outputs are git-ignored and stay in the source tree, where the
Python bindings and the tests load them.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
CSRC = HERE / "csrc"
INCLUDE = REPO / "include"
LIB_OUT = HERE / "qubo_lib_amd" / "libqbx.so"
CPU_OUT = HERE / "qubo_lib_amd" / "libqbx_cpu.so"
EMU_OUT = REPO / "tests" / "_build" / "libqbx_emu.so"
C_CHECK_OUT = REPO / "tests" / "_build" / "qbx_c_check"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 (v_pk_*_f32) stays off: the kernels must not depend on it
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed ({r.returncode}): {' '.join(map(str, cmd))}")
    return r


def build_lib(out: Path = LIB_OUT) -> Path:
    out.parent.mkdir(parents=True, exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-o", str(tmp), str(CSRC / "qbx.hip"),
          str(CSRC / "qbx_server.hip"), "-lpthread"])
    os.replace(tmp, out)
    return out


def build_cpu(out: Path = CPU_OUT) -> Path:
    """The device='cpu' backend: generic x86-64 code (no -march: it runs on any x86-64 host)."""
    out.parent.mkdir(parents=True, exist_ok=True)
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wno-unknown-pragmas",
          f"-I{INCLUDE}", f"-I{CSRC}", "-o", str(tmp), str(CSRC / "qbx_cpu.cpp")])
    os.replace(tmp, out)
    return out


def build_emu(out: Path = EMU_OUT) -> Path:
    out.parent.mkdir(parents=True, exist_ok=True)
    # no floating-point contraction: the generator's float scaling is one rounded multiply
    # and one rounded add, as numpy evaluates it
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", f"-I{CSRC}", "-o", str(out),
          str(CSRC / "qbx_emu.cpp")])
    return out


def build_c_check(out: Path = C_CHECK_OUT, lib: Path = LIB_OUT) -> Path:
    """A plain C99 program linked against libqbx.so (tests/c/qbx_c_check.c): shows that
    include/qbx.h serves C callers as it is."""
    out.parent.mkdir(parents=True, exist_ok=True)
    _run(["gcc", "-std=c99", "-Wall", "-Werror", "-O1", f"-I{INCLUDE}", "-o", str(out),
          str(REPO / "tests" / "c" / "qbx_c_check.c"), f"-L{lib.parent}", "-lqbx",
          "-Wl,-rpath,$ORIGIN/" + os.path.relpath(lib.parent, out.parent)])
    return out


def build_all() -> list[Path]:
    return [build_lib(), build_cpu(), build_emu(), build_c_check()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
