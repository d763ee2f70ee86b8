"""The batched and serving workloads at their own configuration.

  * the 4096-problem uneven mix (MaxCut G(n, 1/2) with n ~ U{16..32} and integer TSP of
    4-5 cities, batch_mix(4096, seed 0)) solved in ONE persistent launch, both as a
    captured batch plan (HIP graph) replayed twice and as an eager batch: every answer
    equals the problem's own solve, and members of n <= 22 equal the oracle's brute force;
  * the serving request stream (serving_requests: Poisson arrivals at 20 000 requests/s,
    n in 20..36 with P(n) ~ 2^-(n-20), half symmetric int16 RandomQUBO in [-127, 127],
    half MaxCut G(n, 1/2)) submitted on its own arrival schedule to the continuous-batching
    server: every (E, state) equals the plain solve's, members of n <= 22 equal the
    oracle, and each request's timestamps are ordered.
These replace the reference's one-solve-per-task loop (src/qubo_lib/solvers/parallel.py:
145-157, 381-385; qbf.py:48-70), whose answers are the individual solves'.
"""
import sys
import time

import numpy as np
import pytest

import oracle
from qubo_lib_amd import engine
from qubo_lib_amd.problems import generators as gen
from qubo_lib_amd.serve import Server

pytestmark = pytest.mark.gpu

ORACLE_MAX_N = 22


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_batch_mix_4096_one_launch(gpu_ctx):
    probs, kinds = gen.batch_mix(4096, seed=0)
    Qs = [p.build_Q_matrix() for p in probs]
    preps = [engine.prepare(q) for q in Qs]
    assert all(p.kind == "i32" for p in preps)
    mats = [p.matrix for p in preps]
    ns = [m.shape[0] for m in mats]
    assert set(ns) <= set(range(16, 33)) and kinds.count("tsp") == 2048
    packed = gpu_ctx.pack(mats, np.int32)
    say("batch_mix(4096, seed 0): graph replays")
    with gpu_ctx.batch_plan(ns, kind="i32") as plan:
        replays = [[(r.E_int, r.state, r.status) for r in plan.run(packed=packed)] for _ in range(2)]
    say("batch_mix(4096, seed 0): eager batch")
    eager = [(r.E_int, r.state, r.status) for r in gpu_ctx.batch(mats, kind="i32", packed=packed)]
    assert replays[0] == replays[1] == eager
    assert all(s == 0 for _, _, s in eager)
    say("batch_mix(4096, seed 0): individual solves and the oracle")
    checked = 0
    for i, (Q, p) in enumerate(zip(Qs, preps)):
        alone = engine.solve_matrix(p)
        assert (alone.E, alone.state) == eager[i][:2], (i, kinds[i], Q.shape)
        if Q.shape[0] <= ORACLE_MAX_N:
            assert eager[i][:2] == oracle.brute_i64(Q), (i, kinds[i], Q.shape)
            checked += 1
    assert checked > 1500        # every TSP(4) and the MaxCut members of n <= 22


def test_serving_stream_on_schedule(gpu_ctx):
    reqs = gen.serving_requests(2000, 20000.0, seed=0)
    preps = [engine.prepare(q) for _, q, _ in reqs]
    arrivals = [t for t, _, _ in reqs]
    say("serving_requests(2000, 20000/s, seed 0): submitting on the arrival schedule")
    with Server(max_n=40) as srv:
        t0 = time.perf_counter()
        ids = []
        for a, p in zip(arrivals, preps):
            while time.perf_counter() - t0 < a:
                pass
            ids.append(srv.submit(p))
        got = {}
        deadline = time.time() + 60.0
        while len(got) < len(ids):
            for r in srv.poll(wait_ms=20):
                got[r["id"]] = r
            assert time.time() < deadline, f"only {len(got)} of {len(ids)} results"
        st = srv.stats()
    assert st["completed"] == len(ids) and st["steps"] >= 2
    say("serving stream: individual solves and the oracle")
    prev_submit = -1.0
    checked = 0
    for (t, Q, kind), p, i in zip(reqs, preps, ids):
        r = got[i]
        assert r["status"] == 0
        assert r["t_submit_ms"] <= r["t_admit_ms"] <= r["t_done_ms"]
        assert r["t_submit_ms"] >= prev_submit       # submitted in arrival order
        prev_submit = r["t_submit_ms"]
        alone = engine.solve_matrix(p)
        assert (r["E"], r["state"]) == (alone.E, alone.state), (kind, Q.shape)
        if Q.shape[0] <= ORACLE_MAX_N:
            assert (r["E"], r["state"]) == oracle.brute_i64(Q), (kind, Q.shape)
            checked += 1
    assert checked > 1500
