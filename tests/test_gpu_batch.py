"""Batched launches give exactly the answers of individual solves."""
import numpy as np
import pytest

import oracle
from helpers import KINDS, random_instance
from qubo_lib_amd import engine
from qubo_lib_amd.problems import generators as gen

pytestmark = pytest.mark.gpu


def test_batch_matches_oracle_mixed_sizes(gpu_ctx):
    rng = np.random.default_rng(42)
    Qs = []
    for i in range(120):
        n = int(rng.integers(1, 21))
        Qs.append(random_instance(rng, n, KINDS[i % len(KINDS)]))
    res = engine.solve_matrices(Qs)
    for Q, r in zip(Qs, res):
        assert (r.E, r.state) == oracle.brute_i64(Q)


def test_batch_mix_equals_individual(gpu_ctx):
    probs, kinds = gen.batch_mix(96, seed=3)
    Qs = [p.build_Q_matrix() for p in probs]
    batched = engine.solve_matrices(Qs)
    for Q, b in zip(Qs, batched):
        one = engine.solve_matrix(Q)
        assert (b.E, b.state) == (one.E, one.state)
    # small members against the oracle too
    for Q, b in zip(Qs, batched):
        if Q.shape[0] <= 18:
            assert (b.E, b.state) == oracle.brute_i64(Q)


def test_batch_mixed_paths_and_variants(gpu_ctx):
    rng = np.random.default_rng(9)
    Qs = [rng.integers(-50, 50, (n, n)) for n in (3, 5, 9, 17)]
    Qs += [rng.standard_normal((n, n)).astype(np.float32) for n in (2, 8, 15)]
    Qs += [rng.integers(-2 ** 35, 2 ** 35, (6, 6))]
    res = engine.solve_matrices(Qs)
    for Q, r in zip(Qs, res):
        if np.issubdtype(Q.dtype, np.integer):
            assert (r.E, r.state) == oracle.brute_i64(Q)
        else:
            E, s = oracle.brute_f64(Q.astype(np.float64))
            assert r.state == s and r.E == pytest.approx(E, rel=1e-9, abs=1e-12)


def test_batch_repeated_calls_reuse_state(gpu_ctx):
    """The per-problem reduction state is reset on device at every launch."""
    rng = np.random.default_rng(10)
    Qs = [rng.integers(-9, 9, (12, 12)) for _ in range(20)]
    first = engine.solve_matrices(Qs)
    second = engine.solve_matrices(Qs[::-1])
    assert [(a.E, a.state) for a in first] == [(b.E, b.state) for b in second[::-1]]
