"""qbx_solve_i32_noalloc (the solve for callers that cannot allocate: its device memory is
the library's fixed scratch array) against qbx_solve_i32, through the C ABI: the same
kernels must give the same answer and result fields, call after call, whatever the size of
the previous call (the array keeps the last call's contents)."""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import KINDS, random_instance
from qubo_lib_amd import engine
from qubo_lib_amd.problems import generators as gen

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "oracle_goldens.json").read_text())

FIELDS = ("E_int", "state", "rank", "states", "tiles", "n", "tie_B", "walk_bits", "inner_bits", "lane_bits",
          "status", "path")


def same(a, b):
    return all(getattr(a, f) == getattr(b, f) for f in FIELDS)


def check(ctx, Q, **kw):
    m = engine.prepare(Q).matrix
    a = ctx.solve_i32_noalloc(m, **kw)
    b = ctx.solve_i32(m, **kw)
    assert same(a, b), (Q.shape, kw, a.as_dict(), b.as_dict())
    return a


def test_noalloc_equals_solve_i32(gpu_ctx):
    print("no-alloc: n=40, then a smaller n=32 in the same scratch", file=sys.stderr, flush=True)
    big = check(gpu_ctx, gen.symmetric_int_qubo(40, 0))
    assert big.path == 3                                  # the rebased pair walk
    assert (big.E_int, big.state) == (GOLD["sym_int16_n40_s0"]["E"], GOLD["sym_int16_n40_s0"]["state"])
    small = check(gpu_ctx, gen.symmetric_int_qubo(32, 0))
    assert (small.E_int, small.state) == (GOLD["sym_int16_n32_s0"]["E"], GOLD["sym_int16_n32_s0"]["state"])
    # the other walks, tie widths and tilings; small problems against the oracle
    check(gpu_ctx, gen.maxcut_gnp(36, 0.5, seed=3))       # packed pair walk
    check(gpu_ctx, gen.tsp_int(5, seed=4))                # int32 walk
    check(gpu_ctx, gen.symmetric_int_qubo(34, 5), tie_B=7, walk_bits=16, inner_bits=6)
    rng = np.random.default_rng(90)
    for t in range(60):
        n = int(rng.integers(1, 21))
        Q = random_instance(rng, n, KINDS[t % len(KINDS)])
        B = int(rng.integers(0, n + 1))
        r = check(gpu_ctx, Q, tie_B=B)
        assert (r.E_int, r.state) == oracle.brute_i64(Q, B)
    check(gpu_ctx, gen.symmetric_int_qubo(43, 6))         # the largest n at the default tiling


def test_noalloc_rejects_what_does_not_fit(gpu_ctx):
    with pytest.raises(ValueError, match="scratch"):
        gpu_ctx.solve_i32_noalloc(np.zeros((44, 44), dtype=np.int32))
    with pytest.raises(OverflowError):
        gpu_ctx.solve_i32_noalloc(np.full((8, 8), 2 ** 28, dtype=np.int32))
    # a rejected call leaves the next one intact
    check(gpu_ctx, gen.symmetric_int_qubo(24, 7))
