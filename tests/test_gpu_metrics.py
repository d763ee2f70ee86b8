"""Per-prefix metrics recorded by the GPU walk: the reference's energy_arr at the n = 40
headline configuration (A = 20: 2^20 prefixes), small problems against enumeration, and
the GPU's arrays equal to the host backend's."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import KINDS, random_instance
from qubo_lib_amd import engine
from qubo_lib_amd.problems import Just_Q
from qubo_lib_amd.problems import generators as gen
from qubo_lib_amd.solvers import QBF
from test_metrics import brute_prefix_minima, oracle_prefix_minima

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "oracle_goldens.json").read_text())


def test_energy_arr_at_n40(gpu_ctx):
    Q = gen.symmetric_int_qubo(40, 0, -127, 127, np.int16)
    print("energy_arr n=40 int16 seed 0 (2^20 prefixes)", file=sys.stderr, flush=True)
    sol = QBF().solve(Just_Q(Q), device="cuda", use_metrics="all")
    arr, tarr = sol.metrics[0]["energy_arr"], sol.metrics[0]["time_arr"]
    g = GOLD["sym_int16_n40_s0"]
    assert arr.shape == (2 ** 20,) and tarr.shape == (2 ** 20,)
    assert sol.best_E == g["E"] and int("".join(map(str, sol.best_X)), 2) == g["state"]
    assert arr.min() == g["E"]
    # the answer's prefix (Gray visiting index of its top 20 bits) holds the minimum
    p = g["state"] >> 20
    i_best = int(oracle.rank_of_state(p, 0))
    assert arr[i_best] == g["E"]
    rng = np.random.default_rng(81)
    idx = sorted({0, 1, 2 ** 20 - 1, i_best, *map(int, rng.integers(0, 2 ** 20, 20))})
    assert np.array_equal(arr[idx], oracle_prefix_minima(Q, 20, idx))
    assert (tarr > 0).all() and sol.run_info["walk"] == "pair16_rebased"


def test_minima_small_problems(gpu_ctx):
    rng = np.random.default_rng(82)
    for t in range(80):
        n = int(rng.integers(1, 13))
        Q = random_instance(rng, n, KINDS[t % len(KINDS)])
        B = int(rng.integers(1, n + 1))
        r, minima, _ = engine.solve_matrix_minima(Q, device=0, tie_B=B)
        assert np.array_equal(minima, brute_prefix_minima(Q, B)), (n, B)
        assert (r.E, r.state) == oracle.brute_i64(Q, B)
    for t in range(30):
        n = int(rng.integers(1, 12))
        Q = rng.standard_normal((n, n)).astype(np.float32)
        B = int(rng.integers(1, n + 1))
        r, minima, _ = engine.solve_matrix_minima(Q, device=0, tie_B=B)
        assert np.allclose(minima, brute_prefix_minima(Q.astype(np.float64), B), rtol=1e-12, atol=1e-12)
        E, s = oracle.brute_f64(Q.astype(np.float64), B)
        assert r.state == s


def test_gpu_minima_equal_host_minima(gpu_ctx):
    cases = [(gen.symmetric_int_qubo(26, 1), 20), (gen.maxcut_gnp(27, 0.5, seed=2), 18),
             (gen.symmetric_int_qubo(24, 2, -100, 100, np.int32), 6),
             (gen.symmetric_normal_qubo(25, 3), 20), (gen.symmetric_normal_qubo(22, 4), 9)]
    for Q, B in cases:
        g, gm, gt = engine.solve_matrix_minima(Q, device=0, tie_B=B)
        c, cm, _ = engine.solve_matrix_minima(Q, device="cpu", tie_B=B)
        # same inner-block width (lanes capped at B on both): FP64 sums in the same order
        assert np.array_equal(gm, cm), (Q.shape, B)
        assert (g.E, g.state) == (c.E, c.state) == (engine.solve_matrix(Q, device=0, tie_B=B).E,
                                                    engine.solve_matrix(Q, device=0, tie_B=B).state)
        assert gm.min() == g.E and gt.shape == gm.shape
