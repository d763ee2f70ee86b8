"""Writes tests/golden/oracle_goldens.json: exact optima of the benchmark workloads,
computed with the CPU oracle's independent brute force (oracle/qbf_oracle.c).

    python tests/golden/make_golden.py [name ...]

Each entry: the workload's generator call, n, tie_B, the minimum energy and the argmin
state (integer, x[0] = most significant bit).  Entries already present are kept; the n=40
ones take several minutes each on 8 cores.
"""
from __future__ import annotations

import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "evo-quantum-qubolib_amd"))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from qubo_lib_amd.problems import RandomQUBO, TSP  # noqa: E402
from qubo_lib_amd.problems.generators import maxcut_gnp, symmetric_int_qubo, symmetric_normal_qubo  # noqa: E402

OUT = Path(__file__).resolve().parent / "oracle_goldens.json"


def workloads():
    w = {}
    for s in (0, 1, 2):
        w[f"sym_int32_n24_s{s}"] = ("symmetric_int_qubo(24, seed, -100, 100, int32)", 24, s,
                                    lambda s=s: symmetric_int_qubo(24, s, -100, 100, np.int32))
        w[f"sym_int16_n32_s{s}"] = ("symmetric_int_qubo(32, seed, -127, 127, int16)", 32, s,
                                    lambda s=s: symmetric_int_qubo(32, s, -127, 127, np.int16))
        w[f"sym_int16_n40_s{s}"] = ("symmetric_int_qubo(40, seed, -127, 127, int16)", 40, s,
                                    lambda s=s: symmetric_int_qubo(40, s, -127, 127, np.int16))
    w["sym_normal_n34_s0"] = ("symmetric_normal_qubo(34, seed)", 34, 0, lambda: symmetric_normal_qubo(34, 0))
    w["sym_normal_n40_s0"] = ("symmetric_normal_qubo(40, seed)", 40, 0, lambda: symmetric_normal_qubo(40, 0))
    w["maxcut_n40_s0"] = ("maxcut_gnp(40, 0.5, seed=seed)", 40, 0, lambda: maxcut_gnp(40, 0.5, seed=0))
    w["tsp6_seed42_f32"] = ("TSP.init_random_instance(seed=42, n_cities=6)", 36, 42,
                            lambda: TSP.init_random_instance(seed=42, n_cities=6).build_Q_matrix())
    w["random_n37_seed42_f32"] = ("RandomQUBO.init_random_instance(seed=42, N=37)", 37, 42,
                                  lambda: RandomQUBO.init_random_instance(seed=42, N=37).build_Q_matrix())
    return w


def main(names):
    data = json.loads(OUT.read_text()) if OUT.exists() else {}
    threads = int(os.environ.get("ORACLE_THREADS", "0"))
    for name, (desc, n, seed, make) in workloads().items():
        if names and name not in names:
            continue
        if name in data:
            continue
        Q = make()
        t0 = time.time()
        if np.issubdtype(Q.dtype, np.integer):
            E, s = oracle.brute_i64(Q, None, threads)
        else:
            E, s = oracle.brute_f64(Q.astype(np.float64), None, threads)
        dt = time.time() - t0
        data[name] = dict(generator=desc, n=n, seed=seed, tie_B=min(n, 20), E=E, state=s,
                          state_bits="".join(str((s >> (n - 1 - i)) & 1) for i in range(n)),
                          oracle="brute", seconds=round(dt, 1))
        OUT.write_text(json.dumps(data, indent=1, sort_keys=True))
        print(name, E, hex(s), f"{dt:.1f}s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
