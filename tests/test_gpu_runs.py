"""Waves that walk several chunks of one problem: problems tiled into more chunks than the
walk grid has waves (each wave takes its first chunk by its workgroup index and the rest
from the ticket counter, and commits its consecutive chunks of a problem once, with the
best block of the run re-walked once).  Integer problems against the host backend (the
same lane walk, reduced on the host: bit-exact) and, up to n = 26, the oracle's brute
force; float problems against the oracle's FP64 brute force.  Random kinds, tie widths and
walk variants; batches whose problems alternate in size, so runs end at problem changes."""
import numpy as np
import pytest

import oracle
from helpers import KINDS, random_instance
from qubo_lib_amd import engine

pytestmark = pytest.mark.gpu

VARIANTS = [(7, 13), (6, 14), (4, 16), (1, 19)]


def _grid_waves(ctx) -> int:
    return int(ctx.device_info()["compute_units"]) * 8     # 2 waves per SIMD, 4 SIMDs per CU


def test_integer_runs_of_chunks(gpu_ctx):
    rng = np.random.default_rng(2028)
    waves = _grid_waves(gpu_ctx)
    for trial in range(24):
        n = int(rng.integers(24, 31))
        Q = random_instance(rng, n, KINDS[trial % len(KINDS)])
        B = int(rng.integers(0, n + 1)) if trial % 3 else None
        L = n - int(rng.integers(18, 21))                  # 2^(n-L) >= 2^18 lanes: >= 4096 tiles
        M, _ = [v for v in VARIANTS if v[0] < L][trial % len([v for v in VARIANTS if v[0] < L])]
        p = engine.prepare(Q)
        got = engine.solve_matrix(p, tie_B=B, walk_bits=L, inner_bits=M)
        assert got.info["tiles"] >= 2 * waves, (trial, n, L, got.info["tiles"])
        host = engine.solve_matrix(p, device="cpu", tie_B=B)
        assert (got.E, got.state) == (host.E, host.state), ("host", trial, n, B, L, M, got.info["walk"])
        if n <= 26:
            assert (got.E, got.state) == oracle.brute_i64(Q, B), ("oracle", trial, n, B, L, M)


def test_float_runs_of_chunks(gpu_ctx):
    rng = np.random.default_rng(2029)
    waves = _grid_waves(gpu_ctx)
    for trial in range(12):
        n = int(rng.integers(23, 25))
        Q = random_instance(rng, n, KINDS[trial % len(KINDS)]).astype(np.float64)
        Q = Q + rng.standard_normal(Q.shape) * (1e-3 if trial % 2 else 1.0)
        B = int(rng.integers(0, n + 1)) if trial % 3 else None
        L = n - 18
        M = 4 if L > 4 else 1
        r = gpu_ctx.solve_f64(Q, tie_B=B, walk_bits=L, inner_bits=M)
        assert int(r.tiles) >= 2 * waves, (trial, n, L, int(r.tiles))
        Ef, sf = oracle.brute_f64(Q, B)
        assert r.state == sf and abs(r.E - Ef) <= 1e-12 * max(1.0, np.abs(Q).sum()), (trial, n, B, L, M)


def test_batches_with_runs(gpu_ctx):
    """Large and small problems in one launch: waves walk runs of a large problem's chunks
    between single chunks of small ones."""
    rng = np.random.default_rng(2030)
    for rep in range(3):
        sizes = [int(rng.choice([14, 18, 26, 28])) for _ in range(40)]
        Qs = [random_instance(rng, n, KINDS[i % len(KINDS)]) for i, n in enumerate(sizes)]
        got = engine.solve_matrices(Qs)
        for Q, r in zip(Qs, got):
            if Q.shape[0] <= 26:
                assert (r.E, r.state) == oracle.brute_i64(Q), (rep, Q.shape)
            else:
                h = engine.solve_matrix(Q, device="cpu")
                assert (r.E, r.state) == (h.E, h.state), (rep, Q.shape)
