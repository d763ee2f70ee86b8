"""Randomised parity sweep: random kinds, sizes, tie widths, inner-block variants and walk
widths through the int32 kernel and the float kernel (called directly, so integer and
perturbed matrices also exercise the quantized filter), and mixed batches, all against
the oracle's brute force.  (A 10 000-check run of the same sweep passed as well.)"""
import numpy as np
import pytest

import oracle
from helpers import KINDS, random_instance
from qubo_lib_amd import engine

pytestmark = pytest.mark.gpu

VARIANTS = [(7, 13), (6, 14), (4, 16), (1, 19)]


def test_random_sweep(gpu_ctx):
    rng = np.random.default_rng(2024)
    for trial in range(800):
        n = int(rng.integers(1, 25))
        Q = random_instance(rng, n, KINDS[trial % len(KINDS)])
        B = int(rng.integers(0, n + 1)) if trial % 3 else None
        cands = [v for v in VARIANTS if v[0] <= n]
        M, OMAX = cands[trial % len(cands)]
        L = int(rng.integers(M, min(n, M + OMAX) + 1))
        ref = oracle.brute_i64(Q, B)
        got = engine.solve_matrix(engine.prepare(Q), tie_B=B, walk_bits=L, inner_bits=M)
        assert (got.E, got.state) == ref, ("int", trial, n, B, L, M)
        Qf = Q.astype(np.float64) + (rng.standard_normal(Q.shape) * 1e-3 if trial % 2 else 0.0)
        Ef, sf = oracle.brute_f64(Qf, B)
        r = gpu_ctx.solve_f64(Qf, tie_B=B, walk_bits=L, inner_bits=M)
        assert r.state == sf and abs(r.E - Ef) <= 1e-12 * max(1.0, np.abs(Qf).sum()), ("f64", trial, n, B, L, M)


def test_random_batches(gpu_ctx):
    rng = np.random.default_rng(2025)
    for rep in range(5):
        Qs = [random_instance(rng, int(rng.integers(1, 21)), KINDS[i % len(KINDS)]) for i in range(200)]
        for Q, r in zip(Qs, engine.solve_matrices(Qs)):
            assert (r.E, r.state) == oracle.brute_i64(Q), (rep, Q.shape)


def test_packed_walks_sweep(gpu_ctx):
    """The three integer walks side by side: int16-range QUBOs of random scale (bounds from
    a few thousand, the plain pair walk, to far beyond the rebased walk's reach, the int32
    walk), problems whose guard fails in some groups, random tie widths, variants and walk
    widths; every answer against the oracle, and every walk taken at least 20 times."""
    from helpers import guard_failing_case
    rng = np.random.default_rng(2026)
    seen = {}
    for trial in range(600):
        n = int(rng.integers(8, 23))
        if trial % 10 == 9:
            n = max(n, 17)
            Q = guard_failing_case(rng, n)
        else:
            scale = int(rng.choice([3, 20, 60, 127, 300, 2000, 30000]))
            U = rng.integers(-scale, scale + 1, (n, n))
            Q = np.triu(U) + np.triu(U, 1).T if trial % 2 else U
        B = int(rng.integers(0, n + 1)) if trial % 3 else None
        cands = [v for v in VARIANTS if v[0] < n]
        M, OMAX = cands[trial % len(cands)]
        L = int(rng.integers(M, min(n, M + OMAX) + 1))
        if trial % 10 == 9:
            L = max(L, min(n, 16))
        got = engine.solve_matrix(Q, tie_B=B, walk_bits=L, inner_bits=M)
        assert (got.E, got.state) == oracle.brute_i64(Q, B), ("packed sweep", trial, n, B, L, M, got.info["walk"])
        seen[got.info["walk"]] = seen.get(got.info["walk"], 0) + 1
    assert all(seen.get(w, 0) >= 20 for w in ("int32", "pair16", "pair16_rebased")), seen


def test_packed_float_filter_sweep(gpu_ctx):
    """The float filters side by side: N(0,1), float32-cast, near-tie (quarter units plus
    1e-9 noise), tiny (1e-150) and one-huge-entry problems, random tie widths, variants and
    walk widths (the packed filter from L = M + 2, the int32 filter at L <= M + 1); every
    answer against the FP64 brute force."""
    rng = np.random.default_rng(2027)
    seen = {}
    for trial in range(150):
        n = int(rng.integers(10, 23))
        Q = rng.standard_normal((n, n))
        kind = trial % 5
        if kind == 1:
            Q = Q.astype(np.float32).astype(np.float64)
        elif kind == 2:
            Q = np.round(Q * 4) / 4 + rng.standard_normal((n, n)) * 1e-9
        elif kind == 3:
            Q = Q * 1e-150
        elif kind == 4:
            Q = Q * 1e-3
            Q[0, n - 1] = 1e6
        cands = [v for v in VARIANTS if v[0] + 1 < n]
        M, OMAX = cands[trial % len(cands)]
        L = int(rng.integers(M, min(n, M + OMAX) + 1))
        B = int(rng.integers(0, n + 1)) if trial % 3 else None
        E, s = oracle.brute_f64(Q, B)
        got = engine.solve_matrix(Q, tie_B=B, walk_bits=L, inner_bits=M, float_filter="packed")
        assert got.state == s and abs(got.E - E) <= 1e-13 * np.abs(Q).sum() + 1e-300, \
            ("float sweep", trial, n, kind, B, L, M, got.info["walk"])
        seen[got.info["walk"]] = seen.get(got.info["walk"], 0) + 1
    assert seen.get("float_pair16", 0) >= 100 and seen.get("float", 0) >= 5, seen
