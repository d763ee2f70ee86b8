"""Continuous-batching server: every answer equals the answer of the same problem solved
alone, whether requests arrive together, staggered, or while large ones are running."""
import time

import numpy as np
import pytest

import oracle
from helpers import KINDS, random_instance
from qubo_lib_amd import engine
from qubo_lib_amd.problems import generators as gen
from qubo_lib_amd.serve import Server

pytestmark = pytest.mark.gpu


def drain(srv, count, timeout_s=60.0):
    got = {}
    t0 = time.time()
    while len(got) < count:
        for r in srv.poll(wait_ms=20):
            got[r["id"]] = r
        assert time.time() - t0 < timeout_s, f"only {len(got)} of {count} results"
    return got


def test_server_matches_individual_solves(gpu_ctx):
    rng = np.random.default_rng(3)
    Qs = [random_instance(rng, int(rng.integers(1, 25)), KINDS[i % len(KINDS)]) for i in range(150)]
    Qs += [rng.standard_normal((n, n)).astype(np.float32) for n in (3, 9, 16, 21)]
    with Server(max_n=40, max_inflight=64) as srv:
        ids = [srv.submit(q) for q in Qs]
        got = drain(srv, len(ids))
    for q, i in zip(Qs, ids):
        alone = engine.solve_matrix(q)
        r = got[i]
        assert r["status"] == 0 and r["state"] == alone.state and r["E"] == alone.E, (q.shape, r)
        if q.shape[0] <= 16 and np.issubdtype(q.dtype, np.integer):
            assert (r["E"], r["state"]) == oracle.brute_i64(q)
        assert r["t_submit_ms"] <= r["t_admit_ms"] <= r["t_done_ms"]


def test_server_requests_arriving_while_large_ones_run(gpu_ctx):
    big = [gen.symmetric_int_qubo(36, s) for s in range(3)]     # ~10 ms of work together
    small = [gen.maxcut_gnp(int(n), 0.5, seed=int(n)) for n in range(16, 26)]
    with Server(max_n=40, max_inflight=32, step_states=2.0 ** 31) as srv:
        ids_big = [srv.submit(q) for q in big]
        time.sleep(0.002)
        ids_small = []
        for q in small:
            ids_small.append(srv.submit(q))
            time.sleep(0.0005)
        got = drain(srv, len(big) + len(small))
        st = srv.stats()
    assert st["steps"] > 3 and st["completed"] == len(big) + len(small)
    for q, i in list(zip(big, ids_big)) + list(zip(small, ids_small)):
        alone = engine.solve_matrix(q)
        assert (got[i]["E"], got[i]["state"]) == (alone.E, alone.state)
    # short requests that arrived while the long ones ran were served alongside them,
    # not held behind them
    done_big = max(got[i]["t_done_ms"] for i in ids_big)
    early = [i for i in ids_small if got[i]["t_submit_ms"] < done_big - 2.0]
    assert early and all(got[i]["t_done_ms"] < done_big for i in early)


def test_server_budget_and_slots_limit_admission(gpu_ctx):
    rng = np.random.default_rng(4)
    Qs = [rng.integers(-50, 50, (22, 22)) for _ in range(40)]
    with Server(max_n=24, max_inflight=4, work_budget=2.0 ** 23) as srv:
        ids = [srv.submit(q) for q in Qs]
        got = drain(srv, len(ids))
    for q, i in zip(Qs, ids):
        assert (got[i]["E"], got[i]["state"]) == oracle.brute_i64(q)


def test_server_rejects_oversized(gpu_ctx):
    with Server(max_n=20) as srv:
        with pytest.raises(ValueError):
            srv.submit(np.zeros((21, 21), dtype=np.int32))


def test_server_float_filter_cases_and_tie_width(gpu_ctx):
    """Float requests that stress the quantized filter, and explicit tie widths, give the
    answers of plain solves."""
    rng = np.random.default_rng(21)
    Qs = []
    for n in (12, 17, 20):
        Q = rng.standard_normal((n, n)) * 1e-3
        Q[0, n - 1] = 1e9                      # small terms quantize to zero
        Qs.append(Q)
        Qs.append(rng.standard_normal((n, n)) * 1e-200)
    Bs = [None, 0, 5, None, 3, 7]
    with Server(max_n=24, max_inflight=8) as srv:
        ids = [srv.submit(q, tie_B=b) for q, b in zip(Qs, Bs)]
        got = drain(srv, len(ids))
    for q, b, i in zip(Qs, Bs, ids):
        E, s = oracle.brute_f64(q, b)
        r = got[i]
        assert r["status"] == 0 and r["state"] == s, (q.shape, b)
        assert abs(r["E"] - E) <= 1e-13 * np.abs(q).sum() + 1e-300


def test_server_stop_with_requests_pending(gpu_ctx):
    """Closing the server while requests are queued and running finishes them and returns."""
    rng = np.random.default_rng(22)
    srv = Server(max_n=30, max_inflight=4, work_budget=2.0 ** 26)
    t0 = time.time()
    for _ in range(40):
        srv.submit(rng.integers(-50, 50, (26, 26)))
    srv.close()
    assert time.time() - t0 < 60


def test_submit_many_matches_individual_solves(gpu_ctx):
    """A burst submitted in one call per arithmetic path (int and float mixed): ids come
    back in input order, every answer equals a plain solve; a bad request queues nothing."""
    rng = np.random.default_rng(5)
    Qs = [random_instance(rng, int(rng.integers(1, 25)), KINDS[i % len(KINDS)]) for i in range(300)]
    Qs.insert(7, rng.standard_normal((12, 12)))
    Qs.insert(100, rng.standard_normal((19, 19)).astype(np.float32))
    with Server(max_n=30, max_inflight=64) as srv:
        ids = srv.submit_many(Qs)
        assert len(set(ids)) == len(Qs)
        int_ids = [i for q, i in zip(Qs, ids) if np.issubdtype(q.dtype, np.integer)]
        assert int_ids == sorted(int_ids)
        got = drain(srv, len(ids))
        with pytest.raises(ValueError):
            srv.submit_many([random_instance(rng, 10, "uniform"), random_instance(rng, 31, "uniform")])
        assert srv.stats()["submitted"] == len(Qs)
    for q, i in zip(Qs, ids):
        alone = engine.solve_matrix(q)
        assert got[i]["status"] == 0 and (got[i]["E"], got[i]["state"]) == (alone.E, alone.state), q.shape


def test_server_packed_float_filter_requests(gpu_ctx):
    """Float requests of n = 22 .. 26 run on the packed float filter in the server (its
    batch rule); answers equal the same problems solved alone (the int32 filter below
    n = 28) and, up to n = 24, the FP64 brute force."""
    rng = np.random.default_rng(33)
    Qs = [rng.standard_normal((n, n)) for n in (22, 23, 24, 26, 22, 25)]
    with Server(max_n=40, max_inflight=16) as srv:
        ids = [srv.submit(q) for q in Qs]
        got = drain(srv, len(ids))
    for q, i in zip(Qs, ids):
        alone = engine.solve_matrix(q)
        r = got[i]
        assert alone.info["walk"] == "float"
        assert r["status"] == 0 and (r["E"], r["state"]) == (alone.E, alone.state), q.shape
        if q.shape[0] <= 24:
            E, s = oracle.brute_f64(q)
            assert r["state"] == s and abs(r["E"] - E) <= 1e-13 * np.abs(q).sum()


def test_server_packed_burst_matches_individual_solves(gpu_ctx):
    """submit_packed: a burst the client packed itself (no per-request Python work) gives
    every request its own solve's answer, ids in order.  A second burst and single submits
    on the same server (request records and staging buffers reused) do too."""
    from qubo_lib_amd import _native
    with Server(max_n=40) as srv:
        for seed, count in ((5, 600), (6, 900)):
            reqs = gen.serving_requests(count, 20000.0, seed=seed)
            preps = [engine.prepare(q) for _, q, _ in reqs]
            packed = _native.Context.pack([p.matrix for p in preps], np.int32)
            ids = list(srv.submit_packed(*packed))
            assert ids == sorted(ids)
            singles = preps[::97]
            ids += [srv.submit(p) for p in singles]
            got = drain(srv, len(ids))
            for p, i in zip(preps + singles, ids):
                alone = engine.solve_matrix(p)
                assert (got[int(i)]["E"], got[int(i)]["state"]) == (alone.E, alone.state), (seed, i)
                assert isinstance(got[int(i)]["E"], int)
