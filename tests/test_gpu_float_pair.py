"""The packed float filter on the GPU: float problems walk coarse 16-bit block-relative
energies over their quantized tables (result path "float_pair16") and re-walk the blocks
that pass in FP64.  Answers must be the FP64 brute force's (state exactly, energy to FP64
summation order), for every inner width and walk width; at sizes beyond the brute force,
quarter-integer problems must give the integer walk's answer on four times the matrix
(energies are exact in FP64 there, so E and the tie-ruled state must agree exactly).
"""
import numpy as np
import pytest

import oracle
from helpers import random_instance
from qubo_lib_amd import engine

pytestmark = pytest.mark.gpu


def flush(*a):
    print(*a, flush=True)


def close(E, ref, Q):
    return abs(E - ref) <= 1e-13 * np.abs(Q).sum() + 1e-300


@pytest.mark.parametrize("inner", [7, 6, 4, 1])
def test_packed_float_filter_vs_oracle(gpu_ctx, inner):
    rng = np.random.default_rng(970 + inner)
    omax = {7: 13, 6: 14, 4: 16, 1: 19}[inner]
    for trial in range(8):
        n = int(rng.integers(max(inner + 3, 14), 23))
        Q = rng.standard_normal((n, n))
        if trial % 4 == 1:
            Q = Q.astype(np.float32).astype(np.float64)
        if trial % 4 == 2:      # near-ties: quarter units plus a 1e-9 perturbation
            Q = np.round(Q * 4) / 4 + rng.standard_normal((n, n)) * 1e-9
        B = [None, 0, int(rng.integers(0, n + 1))][trial % 3]
        E_ref, s_ref = oracle.brute_f64(Q, B)
        for L in sorted({inner + 2, min(n, inner + 5), min(n, inner + omax)}):
            flush("float pair", inner, trial, n, B, L)
            got = engine.solve_matrix(Q, tie_B=B, walk_bits=L, inner_bits=inner, float_filter="packed")
            assert got.info["walk"] == "float_pair16", (inner, trial, n, L, got.info["walk"])
            assert got.state == s_ref and close(got.E, E_ref, Q), (inner, trial, n, B, L)


def test_packed_float_filter_matches_int_walk_large(gpu_ctx):
    """n = 30 .. 34: Q/4 of int16-range, ternary and MaxCut QUBOs on the float path against
    Q on the integer walk: same state, E exactly a quarter."""
    rng = np.random.default_rng(971)
    for n, kind in ((30, "ternary"), (32, "maxcut"), (34, "int16")):
        if kind == "int16":
            U = rng.integers(-127, 128, (n, n))
            Q = np.triu(U) + np.triu(U, 1).T
        else:
            Q = random_instance(rng, n, kind)
        for B in (None, 0):
            flush("float pair large", n, kind, B)
            a = engine.solve_matrix(engine.prepare(Q.astype(np.float64) / 4, lattice=False), tie_B=B,
                                    float_filter="packed")
            b = engine.solve_matrix(Q, tie_B=B)
            assert a.info["walk"] == "float_pair16", a.info["walk"]
            assert (a.E * 4, a.state) == (b.E, b.state), (n, kind, B)


def test_packed_float_filter_in_batches_and_shares(gpu_ctx):
    rng = np.random.default_rng(972)
    Qs = [rng.standard_normal((n, n)) for n in rng.integers(10, 21, 30)]
    got = engine.solve_matrices(Qs, float_filter="packed")
    for i, (Q, r) in enumerate(zip(Qs, got)):
        E_ref, s_ref = oracle.brute_f64(Q)
        assert r.state == s_ref and close(r.E, E_ref, Q), i
        packable = r.info["walk_bits"] > r.info["inner_bits"] + 1     # a pair bit and a walked bit
        assert r.info["walk"] == ("float_pair16" if packable else "float"), i
    Q = rng.standard_normal((26, 26))
    full = engine.solve_matrix(Q, walk_bits=14, float_filter="packed")
    assert full.info["walk"] == "float_pair16"
    merged = engine.merge_parts([engine.solve_part(Q, k, 8, walk_bits=14, float_filter="packed")
                                 for k in range(8)])
    assert (merged.E, merged.state) == (full.E, full.state)


def test_float_filter_choice(gpu_ctx):
    """By default the packed filter runs from n = 28 alone and n = 22 in batches; either
    filter can be forced; all give the same answer."""
    rng = np.random.default_rng(973)
    small, large = rng.standard_normal((24, 24)), rng.standard_normal((30, 30))
    assert engine.solve_matrix(small).info["walk"] == "float"
    assert engine.solve_matrix(large).info["walk"] == "float_pair16"
    batch = engine.solve_matrices([small, rng.standard_normal((18, 18))])
    assert [r.info["walk"] for r in batch] == ["float_pair16", "float"]
    for Q in (small, large):
        runs = [engine.solve_matrix(Q, float_filter=f) for f in (None, "packed", "int32")]
        assert [r.info["walk"] for r in runs[1:]] == ["float_pair16", "float"]
        assert len({(r.E, r.state) for r in runs}) == 1
    with pytest.raises(ValueError):
        engine.solve_matrix(small, float_filter="fp16")


@pytest.mark.parametrize("flt", ["packed", "int32"])
def test_lane_end_flush_with_many_records_and_exact_ties(gpu_ctx, flt):
    """The lane-end flush, which the wave does together over every lane's records, under
    load: integer structure (many blocks within the filter's margin of the optimum) with
    a tiny perturbation, and integer-valued matrices beyond the int32 bound, whose FP64
    energies are exact, so that equal minima in different lanes are decided by the rank.
    Answers equal the FP64 brute force and the host build's lane-by-lane flush."""
    rng = np.random.default_rng(4242)
    for trial in range(10):
        n = int(rng.integers(14, 22))
        base = rng.integers(-2, 3, (n, n)).astype(np.float64)
        if trial % 2:
            Q = base * 2.0 ** 33                       # exact FP64 ties, f64 path
        else:
            Q = base + rng.standard_normal((n, n)) * 1e-8
        p = engine.prepare(Q, lattice=False)
        assert p.kind == "f64"
        B = [None, 0, n][trial % 3]
        E_ref, s_ref = oracle.brute_f64(Q, B)
        host = engine.solve_matrix(p, device="cpu", tie_B=B, float_filter=flt)
        assert host.state == s_ref and close(host.E, E_ref, Q), (trial, n, B)
        for L in sorted({min(n, 9), min(n, 12), n}):
            flush("flush", flt, trial, n, B, L)
            got = engine.solve_matrix(p, tie_B=B, walk_bits=L, inner_bits=7, float_filter=flt)
            assert got.state == s_ref and close(got.E, E_ref, Q), (flt, trial, n, B, L)
            assert (got.E, got.state) == (host.E, host.state) or trial % 2 == 0, (flt, trial, n, B, L)
