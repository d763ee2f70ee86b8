"""One share of a solve (qbx_solve_part_*): the part a single GPU would walk when one problem
is spread over several GPUs.  Merging the shares' (E, rank) minima must give exactly the
whole solve's answer, ties included, for the integer and the float path; the shares must
cover every state once.
"""
import numpy as np
import pytest

import oracle
from helpers import KINDS, random_instance
from qubo_lib_amd import engine

pytestmark = pytest.mark.gpu


def flush(*a):
    print(*a, flush=True)


def shares(Q, parts, **kw):
    return [engine.solve_part(Q, k, parts, **kw) for k in range(parts)]


@pytest.mark.parametrize("kind", KINDS)
def test_int_shares_merge_to_oracle(gpu_ctx, kind):
    rng = np.random.default_rng(31)
    for trial, parts in enumerate((2, 3, 7, 16)):
        n = 22
        Q = random_instance(rng, n, kind)
        B = [None, 0, 9, n][trial]
        flush("int shares", kind, n, parts, B)
        ref = oracle.brute_i64(Q, B)
        res = shares(Q, parts, tie_B=B, walk_bits=7, inner_bits=7)
        merged = engine.merge_parts(res)
        assert (merged.E, merged.state) == ref, (kind, parts, B)
        assert merged.info["states"] == 2 ** n
        assert sum(r.info["tiles"] for r in res) == 2 ** (n - 7 - 6)


def test_int_shares_n30_match_full_solve(gpu_ctx):
    rng = np.random.default_rng(32)
    n = 30
    for kind in ("ternary", "maxcut", "zero"):
        Q = random_instance(rng, n, kind)
        full = engine.solve_matrix(Q, walk_bits=16)
        for parts in (8, 256):
            flush("int n30 shares", kind, parts)
            merged = engine.merge_parts(shares(Q, parts, walk_bits=16))
            assert (merged.E, merged.state, merged.rank) == (full.E, full.state, full.rank), (kind, parts)
            assert merged.info["states"] == 2 ** n


def test_float_shares_match_full_solve(gpu_ctx):
    rng = np.random.default_rng(33)
    for n, parts in ((20, 4), (26, 8)):
        Q = rng.standard_normal((n, n))
        flush("float shares", n, parts)
        full = engine.solve_matrix(Q, walk_bits=12)
        merged = engine.merge_parts(shares(Q, parts, walk_bits=12))
        # every share re-walks its candidate blocks in FP64 from scratch: the same state
        # gets the same energy, bit for bit, whichever launch found it
        assert (merged.E, merged.state) == (full.E, full.state), (n, parts)
        if n == 20:
            E_ref, s_ref = oracle.brute_f64(Q, None)
            assert merged.state == s_ref
            assert abs(merged.E - E_ref) <= 1e-9 * max(1.0, abs(E_ref))


def test_share_arguments_checked(gpu_ctx):
    Q = np.ones((22, 22), dtype=np.int64)
    tiles = 2 ** (22 - 7 - 6)
    for part, parts in ((0, 0), (-1, 2), (2, 2), (0, tiles + 1)):
        with pytest.raises(ValueError):
            engine.solve_part(Q, part, parts, walk_bits=7, inner_bits=7)
    one = engine.solve_part(Q, 0, 1)
    full = engine.solve_matrix(Q)
    assert (one.E, one.state) == (full.E, full.state)
