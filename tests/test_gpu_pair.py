"""The packed pair walk on the GPU: integer problems with energy bound <= 15359 walk two
states per 32-bit operation (16-bit biased energies, per-half minima by
v_pk_minimum3_f16).  Answers must be the oracle's bit for bit, at both ends of the 16-bit
range, for every kernel variant and walk width, and equal to the int32 walk's on the same
problem scaled out of the pair range.
"""
import numpy as np
import pytest

import oracle
from helpers import PAIR_BOUND, energy_bound, pair_case, random_instance
from qubo_lib_amd import engine

pytestmark = pytest.mark.gpu


def flush(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("inner", [7, 6, 4, 1])
def test_pair_walk_vs_oracle(gpu_ctx, inner):
    rng = np.random.default_rng(900 + inner)
    omax = {7: 13, 6: 14, 4: 16, 1: 19}[inner]
    for trial in range(15):
        n = int(rng.integers(max(inner + 1, 10), 23))
        Q = pair_case(rng, n, trial)
        assert energy_bound(Q) <= PAIR_BOUND
        B = [None, 0, int(rng.integers(0, n + 1))][trial % 3]
        ref = oracle.brute_i64(Q, B)
        for L in sorted({inner + 1, min(n, inner + 5), min(n, inner + omax)}):
            flush("pair", inner, trial, n, B, L)
            got = engine.solve_matrix(Q, tie_B=B, walk_bits=L, inner_bits=inner)
            assert (got.E, got.state) == ref, (inner, trial, n, B, L)
            assert got.info["walk"] == "pair16"


def test_pair_range_edges(gpu_ctx):
    """Energies at exactly -15359 and +15359, the bound just above the pair range (another
    walk), and no outer bit to pair on (int32 walk)."""
    n = 20
    rng = np.random.default_rng(5)
    low = -rng.integers(0, 40, (n, n))
    low[0, 0] -= PAIR_BOUND - energy_bound(low)
    high = rng.integers(0, 40, (n, n))
    high[0, 0] += PAIR_BOUND - energy_bound(high)
    for Q, walk in ((low, "pair16"), (high, "pair16"), (-high, "pair16")):
        assert energy_bound(Q) == PAIR_BOUND
        got = engine.solve_matrix(Q, walk_bits=12)
        assert (got.E, got.state) == oracle.brute_i64(Q), walk
        assert got.info["walk"] == walk
    over = low.copy()
    over[0, 0] -= 1
    got = engine.solve_matrix(over, walk_bits=12)
    assert got.info["walk"] != "pair16" and (got.E, got.state) == oracle.brute_i64(over)
    Q = random_instance(rng, 12, "ternary")
    got = engine.solve_matrix(Q, walk_bits=7, inner_bits=7)
    assert got.info["walk"] == "int32" and (got.E, got.state) == oracle.brute_i64(Q)


def test_pair_matches_int32_walk_large(gpu_ctx):
    """n = 30 .. 34 MaxCut and ternary problems: the pair walk's answer equals the int32
    walk's on the same problem scaled by k = 4096 (same argmin, energies x k; couplings too
    large for either packed walk)."""
    rng = np.random.default_rng(6)
    for n, kind in ((30, "maxcut"), (32, "ternary"), (34, "maxcut")):
        Q = random_instance(rng, n, kind)
        k = 4096
        for B in (None, 0):
            flush("pair large", n, kind, B, k)
            a = engine.solve_matrix(Q, tie_B=B)
            b = engine.solve_matrix(Q * k, tie_B=B)
            assert a.info["walk"] == "pair16" and b.info["walk"] == "int32"
            assert (a.E * k, a.state) == (b.E, b.state), (n, kind, B)


def test_pair_in_batches_and_shares(gpu_ctx):
    """Batches mixing pair and int32 problems, and the shares of a pair solve."""
    rng = np.random.default_rng(7)
    Qs = []
    for i in range(40):
        n = int(rng.integers(8, 21))
        Qs.append(pair_case(rng, n, i) if i % 2 else random_instance(rng, n, "uniform") * 200)
    got = engine.solve_matrices(Qs)
    for i, (Q, r) in enumerate(zip(Qs, got)):
        assert (r.E, r.state) == oracle.brute_i64(Q), i
        pair = energy_bound(Q) <= PAIR_BOUND and r.info["walk_bits"] > r.info["inner_bits"]
        assert r.info["walk"] == ("pair16" if pair else "int32"), i
    Q = random_instance(rng, 28, "maxcut")
    full = engine.solve_matrix(Q, walk_bits=16)
    merged = engine.merge_parts([engine.solve_part(Q, k, 8, walk_bits=16) for k in range(8)])
    assert (merged.E, merged.state) == (full.E, full.state)


def sym_int16(rng, n, scale=127):
    U = rng.integers(-scale, scale + 1, (n, n))
    return np.triu(U) + np.triu(U, 1).T


@pytest.mark.parametrize("inner", [7, 6, 4, 1])
def test_rebased_pair_walk_vs_oracle(gpu_ctx, inner):
    """Random int16-range QUBOs (bounds far beyond 15359) take the rebased pair walk:
    16-bit energies relative to each block's base."""
    rng = np.random.default_rng(950 + inner)
    omax = {7: 13, 6: 14, 4: 16, 1: 19}[inner]
    for trial in range(10):
        n = int(rng.integers(max(inner + 2, 16), 23))
        Q = sym_int16(rng, n)
        while energy_bound(Q) <= PAIR_BOUND:     # beyond the plain pair walk's range
            Q = sym_int16(rng, n)
        B = [None, 0, int(rng.integers(0, n + 1))][trial % 3]
        ref = oracle.brute_i64(Q, B)
        for L in sorted({inner + 1, min(n, inner + 5), min(n, inner + omax)}):
            flush("rebased", inner, trial, n, B, L)
            got = engine.solve_matrix(Q, tie_B=B, walk_bits=L, inner_bits=inner)
            assert (got.E, got.state) == ref, (inner, trial, n, B, L)
            assert got.info["walk"] == "pair16_rebased", (inner, trial, n, L)


def test_rebased_pair_guard_failures(gpu_ctx):
    """Problems whose guard fails in 1/16 of the groups (those groups run in int32 on the
    device, wave by wave): answers stay the oracle's."""
    from helpers import guard_failing_case
    rng = np.random.default_rng(951)
    for n in (18, 20, 22):
        Q = guard_failing_case(rng, n)
        for B in (None, 0, 13):
            flush("guard failures", n, B)
            got = engine.solve_matrix(Q, tie_B=B, walk_bits=16)
            assert got.info["walk"] == "pair16_rebased"
            assert (got.E, got.state) == oracle.brute_i64(Q, B), (n, B)


def test_rebased_matches_int32_walk_large(gpu_ctx):
    """n = 30 .. 36 int16-range QUBOs: the rebased pair walk's answer equals the int32
    walk's on the same problem scaled by 3 (same argmin, energies x3; the scaled couplings
    leave the guard no room, so it walks in int32)."""
    rng = np.random.default_rng(952)
    for n in (30, 33, 36):
        Q = sym_int16(rng, n)
        for B in (None, 0):
            flush("rebased large", n, B)
            a = engine.solve_matrix(Q, tie_B=B)
            b = engine.solve_matrix(Q * 3, tie_B=B)
            assert a.info["walk"] == "pair16_rebased" and b.info["walk"] == "int32", (a.info["walk"], b.info["walk"])
            assert (a.E * 3, a.state) == (b.E, b.state), (n, B)
