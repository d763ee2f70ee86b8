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


def test_batch_plan_graph_replays_match_eager(gpu_ctx):
    """A captured batch plan gives the eager batch's answers on every replay, with other
    eager launches on the same stream between the replays, and for new matrices of the
    same shape."""
    probs, _ = gen.batch_mix(64, seed=5)
    mats = [engine.prepare(p.build_Q_matrix()).matrix for p in probs]
    ns = [m.shape[0] for m in mats]
    rng = np.random.default_rng(6)
    other = [rng.integers(-40, 40, (n, n)).astype(np.int32) for n in ns]
    eager = [(r.E_int, r.state) for r in gpu_ctx.batch(mats, kind="i32")]
    eager_other = [(r.E_int, r.state) for r in gpu_ctx.batch(other, kind="i32")]
    with gpu_ctx.batch_plan(ns, kind="i32") as plan:
        for rep in range(4):
            got = [(r.E_int, r.state) for r in plan.run(mats)]
            assert got == eager, f"replay {rep}"
            # eager work of another shape between replays
            gpu_ctx.solve_i32(other[rep])
            got = [(r.E_int, r.state) for r in plan.run(other)]
            assert got == eager_other, f"replay {rep} (second matrix set)"
        with pytest.raises(ValueError):
            plan.run(mats[:-1])
    for Q, e in zip(mats, eager):
        if Q.shape[0] <= 16:
            assert e == oracle.brute_i64(Q)


def test_batch_plan_float(gpu_ctx):
    rng = np.random.default_rng(8)
    Qs = [rng.standard_normal((n, n)) for n in (5, 11, 17, 20, 9)]
    with gpu_ctx.batch_plan([q.shape[0] for q in Qs], kind="f64") as plan:
        for _ in range(2):
            for Q, r in zip(Qs, plan.run(Qs)):
                E, s = oracle.brute_f64(Q)
                assert r.state == s and r.E == pytest.approx(E, rel=1e-12, abs=1e-12)


def test_walk_table_changes_tiling_not_answers(gpu_ctx):
    rng = np.random.default_rng(12)
    Q = rng.integers(-60, 60, (26, 26)).astype(np.int32)
    ref = gpu_ctx.solve_i32(Q)
    table = [-1] * 57
    try:
        gpu_ctx.set_walk_table("i32", "single", [])        # built-in rule only
        base = gpu_ctx.plan_info("i32", "single", 26).walk_bits
        table[26] = base + 2
        gpu_ctx.set_walk_table("i32", "single", table)
        assert gpu_ctx.plan_info("i32", "single", 26).walk_bits == base + 2
        r = gpu_ctx.solve_i32(Q)
        assert r.walk_bits == base + 2
        assert (r.E_int, r.state) == (ref.E_int, ref.state)
        # an entry the variant cannot use falls back to the built-in rule
        table[26] = 1
        gpu_ctx.set_walk_table("i32", "single", table)
        assert gpu_ctx.plan_info("i32", "single", 26).walk_bits == base
        with pytest.raises(ValueError):
            gpu_ctx.set_walk_table("i32", "single", [99] * 57)
    finally:
        gpu_ctx.set_walk_table("i32", "single", [])
        gpu_ctx.load_tuning()


def test_chunk_count_cap_keeps_answers(gpu_ctx):
    """A tiling with more than 2^22 tiles groups several tiles per chunk (the per-chunk
    result slots stay bounded); the answer equals the default tiling's."""
    from qubo_lib_amd.problems.generators import symmetric_int_qubo
    Q = symmetric_int_qubo(36, 5).astype(np.int32)
    ref = gpu_ctx.solve_i32(Q)
    r = gpu_ctx.solve_i32(Q, walk_bits=7)          # 2^29 lanes = 2^23 tiles
    assert r.tiles == 1 << 23
    assert (r.E_int, r.state) == (ref.E_int, ref.state)
