#!/usr/bin/env python3
"""Benchmark: exhaustive QUBO solve throughput on one MI355X.

Default workload (the headline metric): RandomQUBO-style n = 40, Q = triu(U) + triu(U,1)^T
with U ~ integers[-127, 127] int16 (PCG64 seed 0), all 2^40 states enumerated exactly with
int32 accumulation.  One step = one complete solve (prep + walk kernels) of that instance;
Q is staged once per solve (3.2 KB) and the answer (16 B) read back.

    python bench.py [--steps K] [--warmup W] [--workload exhaustive|float|batch|serve|copy]
                    [--no-cpu-baseline]

Other workloads: `float` (n = 34 N(0,1) float32 QUBO, FP64-exact argmin; `--float-n` for
another size), `batch` (4096
problems of uneven n in one launch vs one launch per problem), `serve` (Poisson request
stream through the continuous-batching server vs sequential plain solves), `copy` (a device
copy of known size, to calibrate the FETCH_SIZE / WRITE_SIZE counters).

Prints ONE JSON line.  `value` is states/s for the whole job (one GPU) with Q resident in
device memory (HIP-event time of each solve's kernels); `host_inclusive` gives the wall-clock
rate with the per-solve host copies.  `roofline` prices
the walk kernel against the gfx950 int32 VALU peak (CUs x 4 SIMD x 16 lanes x clock) with
2 algorithmic ops per state (one energy update, one comparison); `cpu_baseline` is the
numpy restatement of the reference QBF algorithm (oracle/qbf_numpy.py) timed on a bounded
sample of the same instance on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "evo-quantum-qubolib_amd"))

import numpy as np  # noqa: E402

OPS_PER_STATE = 2


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(Q, seconds_target: float = 12.0) -> dict:
    """Numpy restatement of QBF on the first K prefix steps of the same instance."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle.qbf_numpy import QBFNumpy
    q = QBFNumpy(Q)
    t0 = time.perf_counter()
    q.build_cache()
    build_s = time.perf_counter() - t0
    # calibrate on 16 steps, then run a sample of about `seconds_target`
    _, _, done, sec = q.run(steps=16)
    per_step = sec / max(done, 1)
    steps = int(max(16, min(2 ** q.A, seconds_target / max(per_step, 1e-6))))
    _, _, done, sec = q.run(steps=steps)
    states = done * 2 ** q.B
    return {
        "value": states / sec,
        "unit": "states/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"numpy QBF restatement (oracle/qbf_numpy.py): first {done} of 2^{q.A} Gray prefix steps "
                   f"of the same n={Q.shape[0]} instance, {states:.3e} states in {sec:.2f} s; "
                   f"cache build ({build_s:.1f} s) excluded; host cores available: {os.cpu_count()}"),
    }


def _parallel_worker(job):
    """One Parallel-style partition: the top m bits fixed to `prefix`, the reduced problem
    solved by the numpy QBF restatement for a bounded number of prefix steps."""
    Q, m, prefix, seconds_target = job
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle.qbf_numpy import QBFNumpy
    n = Q.shape[0]
    x = np.array([(prefix >> (m - 1 - i)) & 1 for i in range(m)], dtype=np.int64)
    Q = Q.astype(np.int64)
    R = Q[m:, m:].copy()
    R[np.diag_indices(n - m)] += x @ Q[:m, m:] + Q[m:, :m] @ x      # fixed-bit fields
    q = QBFNumpy(R)
    q.build_cache()
    _, _, done, sec = q.run(steps=16)
    steps = int(max(16, min(2 ** q.A, seconds_target / max(sec / max(done, 1), 1e-6))))
    _, _, done, sec = q.run(steps=steps)
    return done * 2 ** q.B, sec


def cpu_baseline_parallel(Q, seconds_target: float = 6.0) -> dict:
    """The reference's Parallel(devices=['cpu'] * k) split of the same instance: k processes,
    each on a reduced problem with the top log2(k) bits fixed (parallel.py:145-157), all
    running the numpy restatement for a bounded sample; value = total states / slowest."""
    import multiprocessing as mp
    k = 1
    while k * 2 <= min(os.cpu_count() or 1, 16):
        k *= 2
    m = k.bit_length() - 1
    jobs = [(Q, m, p, seconds_target) for p in range(k)]
    with mp.get_context("fork").Pool(k) as pool:
        out = pool.map(_parallel_worker, jobs)
    states = sum(o[0] for o in out)
    sec = max(o[1] for o in out)
    return {"value": states / sec, "unit": "states/s", "cores": k,
            "sample": (f"{k} processes, each the numpy QBF restatement on a reduced n={Q.shape[0] - m} problem "
                       f"(top {m} bits fixed), {states:.3e} states in {sec:.2f} s (slowest process); "
                       f"cache builds excluded")}


def measured_traffic(workload: str):
    """HBM bytes per walk dispatch from the committed PMC summary (profiles/traffic.json):
    (FETCH_SIZE x fetch factor + WRITE_SIZE x write factor) KiB, or None."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    t = json.loads(path.read_text()).get(workload)
    if not t:
        return None
    return (t["fetch_kib"] * t["factor_fetch"] + t["write_kib"] * t["factor_write"]) * 1024.0


def exhaustive(args, ctx, devinfo):
    from qubo_lib_amd import engine
    from qubo_lib_amd.problems.generators import maxcut_gnp, symmetric_int_qubo, symmetric_normal_qubo
    if args.workload == "float":
        n = args.float_n
        Q = symmetric_normal_qubo(n, args.seed)
        workload = f"symmetric N(0,1) float32 QUBO n={n}, seed {args.seed}: filter walk over quantized tables, FP64 exact re-walk"
        dtype = "f32"
    elif args.workload == "maxcut":
        n = 40
        Q = maxcut_gnp(n, 0.5, seed=args.seed)
        workload = (f"MaxCut G(n=40, p=0.5) unit weights (seed {args.seed}), 2^40 states, exact: energy bound "
                    f"<= 15359, packed pair walk (two 16-bit states per 32-bit op)")
        dtype = "int32"
    else:
        n = 40
        Q = symmetric_int_qubo(n, args.seed, -127, 127, np.int16)
        workload = f"RandomQUBO n={n}, symmetric int16 U[-127,127] (seed {args.seed}), 2^{n} states, exact"
        dtype = "int32"
    prep = engine.prepare(Q)
    kw = dict(walk_bits=args.walk_bits, inner_bits=args.inner_bits)
    for _ in range(args.warmup):
        engine.solve_matrix(prep, **kw)
    kms, results = [], []
    # every libqbx call ends with hipStreamSynchronize on its stream, so the clock below
    # brackets completed device work (torch is not loaded: it carries its own HIP runtime)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r = engine.solve_matrix(prep, **kw)
        kms.append(r.info["kernel_ms"])
        results.append((r.E, r.state))
    wall = time.perf_counter() - t0
    states = 2.0 ** n
    ms_per_step = wall * 1000.0 / args.steps
    kernel_ms = float(np.mean(kms))
    # value: Q already resident in HBM -- device time of each solve's kernels (HIP events on
    # the library's stream around prep + walk); the wall-clock rate with the per-solve
    # 3.2 KB host-to-device copy and 32 B result read-back is reported beside it
    value = states * args.steps / (float(np.sum(kms)) * 1e-3)
    value_host = states * args.steps / wall
    peak = devinfo["valu_peak_ops"]
    # the packed walks do two 16-bit operations per lane and instruction: their ceiling is
    # the packed 16-bit VALU rate, twice the int32 one
    packed = r.info["walk"] in ("pair16", "pair16_rebased", "float_pair16")
    if packed:
        peak *= 2
    achieved = states * OPS_PER_STATE / (kernel_ms * 1e-3)
    assert len(set(results)) == 1, "answers differ between steps"
    E, state = results[0]
    parity = {"E": E, "state": state}
    if args.workload == "maxcut":
        # parity against the int32 walk on the same problem scaled out of the packed walks'
        # range (same argmin, energies times k)
        k = 4096   # couplings of 8192 per edge: too large for either packed walk
        ref = engine.solve_matrix(Q * k, **kw)
        parity.update(walk=r.info["walk"], int32_walk={"E": ref.E, "state": ref.state, "scale": k,
                                                       "walk": ref.info["walk"], "kernel_ms": ref.info["kernel_ms"]},
                      matches_int32_walk=bool(ref.E == E * k and ref.state == state))
    gold = ROOT / "tests" / "golden" / "oracle_goldens.json"
    key = {"float": f"sym_normal_n{n}_s{args.seed}", "maxcut": f"maxcut_n40_s{args.seed}"}.get(
        args.workload, f"sym_int16_n40_s{args.seed}")
    if gold.exists() and key:
        g = json.loads(gold.read_text()).get(key)
        if g is not None:
            parity["oracle_state"] = g["state"]
            parity["oracle_E"] = g["E"]
            parity["abs_err_vs_oracle"] = abs(float(g["E"]) - float(E))
            parity["matches_oracle"] = bool(g["state"] == state and
                                            (g["E"] == E if isinstance(E, int) else abs(g["E"] - E) < 1e-9 * max(1, abs(E))))
    line = {
        "metric": {"float": f"states/s at n={n} float32 (FP64-exact argmin)",
                   "maxcut": "states/s at n=40 MaxCut G(40, 0.5) (exact, packed pair walk)"}.get(
                       args.workload, "states/s at n=40 int16 (exact)"),
        "value": value,
        "unit": "states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "int16x2" if packed else dtype,
        "data": "synthetic",
        "config": {"workload": workload, "n": n, "seed": args.seed, "states_per_step": int(states),
                   "walk_bits": r.info["walk_bits"], "inner_bits": r.info["inner_bits"],
                   "lane_bits": r.info["lane_bits"], "tiles": r.info["tiles"], "tie_B": r.tie_B,
                   "walk": r.info["walk"],
                   "kernel_ms_mean": kernel_ms, "tuning": str(ctx.tuning_file.name) if ctx.tuning_file else "built-in rule",
                   "device": devinfo["name"], "compute_units": devinfo["compute_units"],
                   "clock_mhz": devinfo["clock_khz"] / 1000.0},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TOP/s",
                     "frac": achieved / peak,
                     # PMC traffic is committed for the default instances only
                     "traffic": (None if args.workload == "float" and n != 34 else
                                 measured_traffic({"float": "float", "maxcut": "maxcut"}.get(args.workload, "exhaustive"))),
                     "ops_per_state": OPS_PER_STATE,
                     "note": "kernel time = HIP events around prep+walk on the library stream" +
                             (f"; walk {r.info['walk']}: peak = packed 16-bit VALU ops (2 per lane per instruction)"
                              if packed else "")},
        "parity": parity,
        "host_inclusive": {"value": value_host, "unit": "states/s",
                           "note": "wall clock around the K solve calls: staging Q from host memory, "
                                   "launches, result read-back and synchronisation included"},
    }
    if args.shard_parts > 1:
        line["shard"] = shard_shares(prep, args.shard_parts, r, kw)
    if args.cpu_baseline:
        log("timing the CPU baseline ...")
        line["cpu_baseline"] = cpu_baseline(Q)
        if args.workload != "float":
            log("timing the process-parallel CPU baseline ...")
            line["cpu_baseline"]["parallel"] = cpu_baseline_parallel(Q)
    else:
        line["cpu_baseline"] = None
    return line


def shard_shares(prep, parts: int, full, kw) -> dict:
    """One GPU's share of the same solve spread over `parts` GPUs: each share solved once on
    this GPU (qbx_solve_part), with its device time; the merged answer must be the full one."""
    from qubo_lib_amd import engine
    log(f"solving the {parts} shares ...")
    res = [engine.solve_part(prep, k, parts, **kw) for k in range(parts)]
    merged = engine.merge_parts(res)
    ms = [r.info["kernel_ms"] for r in res]
    return {"parts": parts, "share_kernel_ms": ms, "max_share_kernel_ms": max(ms),
            "share_states": [r.info["states"] for r in res],
            "share_rate": {"value": min(r.info["states"] / (r.info["kernel_ms"] * 1e-3) for r in res),
                           "unit": "states/s", "note": "slowest share: its states / its device time"},
            "merged_matches_full": (merged.E, merged.state) == (full.E, full.state)}


def batch(args, ctx, devinfo):
    """4096-problem uneven mix: one persistent launch vs one launch per problem."""
    from qubo_lib_amd import engine
    from qubo_lib_amd.problems.generators import batch_mix
    probs, kinds = batch_mix(args.batch_count, seed=args.seed)
    Qs = [p.build_Q_matrix() for p in probs]
    preps = [engine.prepare(q) for q in Qs]
    mats = [p.matrix for p in preps]
    packed = ctx.pack(mats, np.int32)
    # batched form: the batch shape captured once in a HIP graph, replayed per step
    plan = ctx.batch_plan([m.shape[0] for m in mats], kind="i32")
    for _ in range(args.warmup):
        plan.run(packed=packed)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = plan.run(packed=packed)
    wall_b = (time.perf_counter() - t0) / args.steps
    graph_ms = float(np.mean([res[0].kernel_ms]))
    plan.close()
    # the same launches issued eagerly (no graph)
    for _ in range(args.warmup):
        ctx.batch(mats, kind="i32", packed=packed)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res_e = ctx.batch(mats, kind="i32", packed=packed)
    wall_e = (time.perf_counter() - t0) / args.steps
    # plain form: one C-ABI solve (launch + sync) per problem, same kernels and GPU
    lat = []
    plain = []
    mats_plain = mats if args.batch_plain else []
    t0 = time.perf_counter()
    for m in mats_plain:
        t1 = time.perf_counter()
        plain.append(ctx.solve_i32(m, walk_bits=None))
        lat.append(time.perf_counter() - t1)
    wall_p = time.perf_counter() - t0
    if not args.batch_plain:          # profiling runs: batched launches only
        plain, lat, wall_p = res, [float("nan")], float("nan")
    same = all((a.E_int, a.state) == (b.E_int, b.state) == (e.E_int, e.state)
               for a, b, e in zip(res, plain, res_e))
    states = float(sum(2.0 ** q.shape[0] for q in Qs))
    drawn = device_drawn_random_batch(args, ctx)
    drawn_tsp = device_built_tsp_batch(args, ctx)
    return {
        "metric": "batched solves/s (4096-instance uneven mix)",
        "value": len(Qs) / wall_b,
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall_b * 1000.0,
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"batch_mix({len(Qs)}, seed={args.seed}): MaxCut G(n,0.5) n~U{{16..32}} / int TSP 4-5 cities",
                   "states": states, "form": "one HIP graph replay per step (copy in, prep, persistent walks, copy out)",
                   "tuning": str(ctx.tuning_file.name) if ctx.tuning_file else "built-in rule"},
        "graph_ms": graph_ms,
        "eager": {"solves_per_s": len(Qs) / wall_e, "ms_per_step": wall_e * 1000.0},
        "plain": {"solves_per_s": len(Qs) / wall_p, "wall_s": wall_p, "p50_ms": float(np.percentile(lat, 50) * 1e3),
                  "p99_ms": float(np.percentile(lat, 99) * 1e3)},
        "speedup_vs_plain": wall_p / wall_b,
        "identical_answers": same,
        "random_qubo_drawn_on_device": drawn,
        "tsp_built_on_device": drawn_tsp,
        "states_per_s": states / wall_b,
    }


def device_drawn_random_batch(args, ctx, count: int = 4096) -> dict:
    """RandomQUBO(seed, N ~ U{16..32}, int16 in [-100, 100)) x count, end to end: matrices
    drawn by numpy on the host and staged vs drawn on the device (same bits)."""
    from qubo_lib_amd.problems import RandomQUBO
    rng = np.random.default_rng(args.seed)
    seeds = [int(s) for s in rng.integers(0, 2 ** 31, count)]
    ns = [int(n) for n in rng.integers(16, 33, count)]
    probs = [RandomQUBO.init_random_instance(seed=s, N=n, dtype=np.int16) for s, n in zip(seeds, ns)]
    ctx.random_qubo_batch(seeds[:64], ns[:64], -100, 100, np.int16)        # warm-up
    t0 = time.perf_counter()
    mats = [p.build_Q_matrix() for p in probs]
    t1 = time.perf_counter()
    host = ctx.batch([m.astype(np.int32) for m in mats], kind="i32")
    t2 = time.perf_counter()
    dev = ctx.random_qubo_batch(seeds, ns, -100, 100, np.int16)
    t3 = time.perf_counter()
    same = all((a.E_int, a.state) == (b.E_int, b.state) for a, b in zip(host, dev))
    return {"workload": f"{count} RandomQUBO int16 [-100,100), N ~ U{{16..32}}",
            "host_drawn": {"solves_per_s": count / (t2 - t0), "numpy_ms": (t1 - t0) * 1e3,
                           "solve_ms": (t2 - t1) * 1e3},
            "device_drawn": {"solves_per_s": count / (t3 - t2), "ms": (t3 - t2) * 1e3,
                             "kernel_ms": dev[0].kernel_ms},
            "identical_answers": same}


def device_built_tsp_batch(args, ctx, count: int = 4096) -> dict:
    """TSP.init_random_instance(seed, 4 or 5 cities, [0, 100), float32) x count, end to end:
    cities drawn and QUBOs built on the host and staged vs both done on the device (same
    bits).  Timed from the seeds: the host leg includes drawing the cities."""
    from qubo_lib_amd.problems import TSP
    rng = np.random.default_rng(args.seed + 1)
    seeds = [int(s) for s in rng.integers(0, 2 ** 31, count)]
    ncs = [int(c) for c in rng.integers(4, 6, count)]
    ctx.random_tsp_batch(seeds[:64], ncs[:64], 0, 100, np.float32)          # warm-up
    t0 = time.perf_counter()
    mats = [TSP.init_random_instance(seed=s, n_cities=c).build_Q_matrix() for s, c in zip(seeds, ncs)]
    t1 = time.perf_counter()
    host = ctx.batch([m.astype(np.float64) for m in mats], kind="f64")
    t2 = time.perf_counter()
    dev = ctx.random_tsp_batch(seeds, ncs, 0, 100, np.float32)
    t3 = time.perf_counter()
    same = all((a.E, a.state) == (b.E, b.state) for a, b in zip(host, dev))
    return {"workload": f"{count} TSP float32 cities in [0,100)^2, 4 or 5 cities (N = 16 or 25)",
            "host_built": {"solves_per_s": count / (t2 - t0), "build_ms": (t1 - t0) * 1e3,
                           "solve_ms": (t2 - t1) * 1e3},
            "device_built": {"solves_per_s": count / (t3 - t2), "ms": (t3 - t2) * 1e3,
                             "kernel_ms": dev[0].kernel_ms},
            "identical_answers": same}


def serve(args, ctx, devinfo):
    """Poisson request stream (n in 20..36, P(n) ~ 2^-(n-20); half symmetric int16
    RandomQUBO in [-127,127], half MaxCut G(n,0.5)) served by the continuous-batching
    server at several arrival rates, and by sequential plain solves at the headline rate."""
    from qubo_lib_amd import engine
    from qubo_lib_amd.problems.generators import serving_requests
    from qubo_lib_amd.serve import Server
    count = args.serve_count
    base_rate = 20000.0
    reqs = serving_requests(count, base_rate, seed=args.seed)
    preps = [engine.prepare(q) for _, q, _ in reqs]
    arrivals = np.array([t for t, _, _ in reqs])
    rows = []
    answers = {}
    for rate in args.serve_rates:
        arr = arrivals * (base_rate / rate)
        with Server(max_n=40) as srv:
            # warm-up: kernels loaded, slots touched
            for p in preps[:64]:
                srv.submit(p)
            w = 0
            while w < 64:
                w += len(srv.poll(wait_ms=20))
            t0 = time.perf_counter()
            ids = []
            for a, p in zip(arr, preps):
                while time.perf_counter() - t0 < a:
                    pass
                ids.append(srv.submit(p))
            got = {}
            while len(got) < len(ids):
                for r in srv.poll(wait_ms=20):
                    got[r["id"]] = r
            wall = time.perf_counter() - t0
            st = srv.stats()
        lat = np.array([got[i]["t_done_ms"] - got[i]["t_submit_ms"] for i in ids])
        for j, i in enumerate(ids):
            answers.setdefault(j, set()).add((got[i]["E"], got[i]["state"]))
        rows.append({"rate_per_s": rate, "solves_per_s": len(ids) / wall, "p50_ms": float(np.percentile(lat, 50)),
                     "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "wall_s": wall,
                     "steps": int(st["steps"]), "device_busy_ms": st["busy_ms"]})
        log(f"serve rate={rate:.0f}/s -> {rows[-1]}")
    # capacity: every request submitted at once, in one submit_many call
    with Server(max_n=40) as srv:
        for p in preps[:64]:
            srv.submit(p)
        w = 0
        while w < 64:
            w += len(srv.poll(wait_ms=20))
        t0 = time.perf_counter()
        ids = srv.submit_many(preps)
        t_sub = time.perf_counter() - t0
        got = {}
        while len(got) < len(ids):
            for r in srv.poll(wait_ms=20):
                got[r["id"]] = r
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat = np.array([got[i]["t_done_ms"] - got[i]["t_submit_ms"] for i in ids])
    for j, i in enumerate(ids):
        answers.setdefault(j, set()).add((got[i]["E"], got[i]["state"]))
    burst = {"solves_per_s": len(ids) / wall, "submit_us_per_request": t_sub / len(ids) * 1e6,
             "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
             "steps": int(st["steps"]), "device_busy_ms": st["busy_ms"]}
    log(f"serve burst -> {burst}")
    # sequential plain solves (one blocking launch each) on the same arrivals at the base rate
    t0 = time.perf_counter()
    lat = []
    for j, (a, p) in enumerate(zip(arrivals, preps)):
        while time.perf_counter() - t0 < a:
            pass
        r = engine.solve_matrix(p)
        lat.append((time.perf_counter() - t0 - a) * 1e3)
        answers.setdefault(j, set()).add((r.E, r.state))
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    seq = {"rate_per_s": base_rate, "solves_per_s": len(lat) / wall, "p50_ms": float(np.percentile(lat, 50)),
           "p99_ms": float(np.percentile(lat, 99)), "wall_s": wall}
    same = all(len(v) == 1 for v in answers.values())
    head = max(rows, key=lambda r: r["solves_per_s"] if r["p99_ms"] < 5.0 else 0.0)
    return {
        "metric": "serving solves/s (Poisson arrivals, n 20..36) with p50/p99 latency",
        "value": head["solves_per_s"], "unit": "solves/s",
        "n_gpus": 1, "steps": count, "warmup": 64, "ms_per_step": 1000.0 / head["solves_per_s"],
        "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"serving_requests({count}, seed={args.seed}): n in 20..36, P(n)~2^-(n-20), "
                               "half int16 RandomQUBO [-127,127], half MaxCut G(n,0.5)",
                   "headline_rate_per_s": head["rate_per_s"]},
        "p50_ms": head["p50_ms"], "p99_ms": head["p99_ms"],
        "rates": rows, "burst": burst, "sequential": seq,
        "identical_answers": same,
    }


def copy_calibration(args):
    """A plain device-to-device copy of known size (hipMemcpy -> the runtime's copy
    kernel), used to scale the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so.7")
    nbytes = args.copy_mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    wall = (time.perf_counter() - t0) / args.steps
    hip.hipFree(a)
    hip.hipFree(b)
    return {"metric": "device copy bandwidth (calibration)", "value": 2 * nbytes / wall / 1e9, "unit": "GB/s",
            "n_gpus": 1, "steps": args.steps, "warmup": 0, "ms_per_step": wall * 1e3, "higher_is_better": True,
            "scaling": "strong", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
            "config": {"workload": f"hipMemcpy D2D {args.copy_mib} MiB", "bytes_read": nbytes, "bytes_written": nbytes}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["exhaustive", "float", "maxcut", "batch", "serve", "copy"],
                    default="exhaustive")
    ap.add_argument("--serve-count", type=int, default=20000)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[2000.0, 20000.0, 100000.0])
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--float-n", type=int, default=34, help="size of the float workload's instance")
    ap.add_argument("--walk-bits", type=int, default=None)
    ap.add_argument("--inner-bits", type=int, default=None)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--no-batch-plain", dest="batch_plain", action="store_false",
                    help="batch workload: skip the one-launch-per-problem comparison (for profiling)")
    ap.add_argument("--no-cpu-baseline", dest="cpu_baseline", action="store_false")
    ap.add_argument("--shard-parts", type=int, default=8,
                    help="exhaustive workloads: also solve the problem as this many shares (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(one GPU's share of a sharded solve is measured by --shard-parts; the shares' "
              "16-byte (E, rank) minima merge on the host)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "copy":
        print(json.dumps(copy_calibration(args)), flush=True)
        return
    from qubo_lib_amd import _native
    ctx = _native.get_context(0)
    devinfo = ctx.device_info()
    if args.workload == "batch":
        line = batch(args, ctx, devinfo)
    elif args.workload == "serve":
        line = serve(args, ctx, devinfo)
    else:
        line = exhaustive(args, ctx, devinfo)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
