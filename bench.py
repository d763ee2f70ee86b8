#!/usr/bin/env python3
"""Benchmark: exhaustive QUBO solving on one MI355X.

Headline (the default run): RandomQUBO-style n = 40, Q = triu(U) + triu(U,1)^T with
U ~ integers[-127, 127] int16 (PCG64 seed 0), all 2^40 states enumerated exactly.  One
step = one complete solve (prep + walk kernels) of that instance; Q is staged once per
solve (3.2 KB) and the answer (16 B) read back.  The same JSON line also carries the
metric's second clause and the other configurations, each timed in this run:

  batch     batch_mix(4096, seed 0) -- MaxCut G(n, 1/2), n ~ U{16..32}, and integer TSP of
            4-5 cities -- in one persistent launch (a captured HIP graph, replayed per
            step) against one launch + sync per problem: solves/s, p50/p99 per-problem
            latency of both forms, speedup, identical answers
  serve     serving_requests(20000, seed 0) at 20 000 requests/s through the
            continuous-batching server, against sequential plain solves of the same
            arrivals: solves/s, p50/p99 latency, identical answers
  cpu_c1    the reference's CPU configuration: QBF (numpy restatement) on n = 24 int32
            RandomQUBO-style instances, seeds 0-2, end to end, exact against the goldens;
            beside it the product's own device='cpu' backend on the same instances
  float34   n = 34 float32 N(0,1) instance, FP64-exact argmin: states/s, |E - E_fp64|,
            argmin agreement with the FP64 brute-force golden

    python bench.py [--steps K] [--warmup W] [--workload exhaustive|float|maxcut|batch|serve|copy|float-ties]
                    [--headline-only] [--no-cpu-baseline]

--workload float-ties times both float filters on inputs whose ties FP64 keeps only to
rounding (MaxCut QUBOs scaled by 0.1 and 1/3, a float32 TSP), one row per instance.

`value` is states/s for the whole job (one GPU) with Q resident in device memory (HIP-event
time of each solve's kernels); `host_inclusive` gives the wall-clock rate with the
per-solve host copies.  `roofline` prices the walk against the gfx950 VALU issue rate
(CUs x 4 SIMD x 16 lanes x clock lane-instructions per second): `frac` = states/s x the
fewest VALU instructions per state the walk that ran can issue, / that peak.  The PMC
instruction count and VALU busy of the same walk come from the committed profiles.
`cpu_baseline` is the numpy restatement of the reference QBF algorithm
(oracle/qbf_numpy.py) timed on a bounded sample of the same instance on this host.
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

# The reference's "O(1) per state" (qbf.py:22): one energy update and one comparison.
ALG_OPS_PER_STATE = 2
PACKED_WALKS = ("pair16", "pair16_rebased", "float_pair16")
METRIC = "states/s at n=40 int16 (exact); batched solves/s + p99 vs per-instance launch"
# host threads for the device='cpu' backend: OMP_NUM_THREADS as given at start, else the
# host's cores (at most 16, the host share of one GPU)
HOST_THREADS = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def isa_min_instr_per_state(walk: str, M: int) -> float:
    """Fewest VALU instructions per state the walk can issue.  The inner block is a binary
    trie over M bits: 2^M - 1 states are their parent plus one field plus one constant
    (one v_add3_u32 each) and the 2^M states fold into the block minimum two at a time
    (one v_min3_i32 per pair).  The packed walks hold two states per 32-bit word, so one
    v_add3_u32 advances two states and one v_pk_minimum3_f16 folds four: half the count."""
    per = ((1 << M) - 1 + (1 << (M - 1))) / float(1 << M)
    return per / 2.0 if walk in PACKED_WALKS else per


def _committed(name: str, key: str):
    path = ROOT / "profiles" / name
    if not path.exists():
        return None
    return json.loads(path.read_text()).get(key)


def measured_traffic(key: str):
    """HBM bytes per walk dispatch from the committed PMC summary (profiles/traffic.json):
    FETCH_SIZE x fetch factor + WRITE_SIZE x write factor (KiB), or None."""
    t = _committed("traffic.json", key)
    if not t:
        return None
    return (t["fetch_kib"] * t["factor_fetch"] + t["write_kib"] * t["factor_write"]) * 1024.0


def roofline(states_per_s: float, walk: str, M: int, devinfo: dict, key: str | None) -> dict:
    """VALU roofline of the walk: achieved = states/s x the walk's ISA-minimal VALU
    instructions per state; peak = the device's VALU issue rate in lane-instructions/s."""
    peak = devinfo["valu_peak_ops"]
    imin = isa_min_instr_per_state(walk, M)
    achieved = states_per_s * imin
    out = {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "T VALU lane-instr/s",
           "frac": achieved / peak,
           "traffic": measured_traffic(key) if key else None,
           "walk": walk, "isa_min_instr_per_state": imin, "ceiling_states_per_s": peak / imin,
           "peak_basis": (f"{devinfo['compute_units']} CUs x 4 SIMD x 16 lanes x {devinfo['clock_khz'] / 1e6:.3f} GHz "
                          "(hipGetDeviceProperties)"),
           "alg_ops_rate": {"value": states_per_s * ALG_OPS_PER_STATE / 1e12, "unit": "T ops/s",
                            "note": "2 algorithmic ops per state (one energy update, one comparison; qbf.py:22); "
                                    "a rate, not a bound: one packed instruction retires several of them"}}
    pmc = _committed("pmc.json", key) if key else None
    if pmc:
        out["pmc"] = {"valu_instr_per_state": pmc["valu_instr_per_state"], "valu_busy": pmc["valu_busy"],
                      "issue_frac": states_per_s * pmc["valu_instr_per_state"] / peak, "source": pmc["source"],
                      "note": "issue_frac = states/s (this run) x PMC VALU instructions per state / peak"}
    return out


# ------------------------------------------------------------------------ CPU baselines
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


def cpu_c1() -> dict:
    """The reference's CPU configuration, end to end: QBF (numpy restatement, B = 20,
    A = 4, qbf.py:188-230, 443-446) on the n = 24 int32 instances of seeds 0-2, exact
    (E, argmin) against the brute-force goldens."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle.qbf_numpy import QBFNumpy
    from qubo_lib_amd.problems.generators import symmetric_int_qubo
    gold = json.loads((ROOT / "tests" / "golden" / "oracle_goldens.json").read_text())
    rows = []
    for seed in (0, 1, 2):
        Q = symmetric_int_qubo(24, seed, -100, 100, np.int32)
        t0 = time.perf_counter()
        q = QBFNumpy(Q)
        q.build_cache()
        E, state, done, _ = q.run()
        wall = time.perf_counter() - t0
        g = gold[f"sym_int32_n24_s{seed}"]
        E = int(round(float(E)))
        rows.append({"seed": seed, "E": E, "state": int(state), "wall_s": wall, "states_per_s": 2.0 ** 24 / wall,
                     "matches_golden": bool(E == g["E"] and int(state) == g["state"])})
    wall = sum(r["wall_s"] for r in rows)
    # the product's own device='cpu' backend (libqbx_cpu.so, OpenMP) on the same instances,
    # through the drop-in solver, end to end; best of three solves per instance (a spinning
    # OpenMP team of one thread per core slows many-fold whenever the host runs anything
    # else: one run measured 7.6 ms per solve against 0.4 ms)
    from qubo_lib_amd import _native
    from qubo_lib_amd.problems import Just_Q
    from qubo_lib_amd.solvers import QBF
    _native.get_backend("cpu").threads = HOST_THREADS
    prod = []
    for seed in (0, 1, 2):
        Q = symmetric_int_qubo(24, seed, -100, 100, np.int32)
        QBF().solve(Just_Q(Q), device="cpu")                # warm-up (thread pool)
        best = None
        for _ in range(3):
            t0 = time.perf_counter()
            sol = QBF().solve(Just_Q(Q), device="cpu")
            wall_p = time.perf_counter() - t0
            if best is None or wall_p < best[0]:
                best = (wall_p, sol)
        wall_p, sol = best
        g = gold[f"sym_int32_n24_s{seed}"]
        prod.append({"seed": seed, "wall_s": wall_p, "walk_ms": sol.kernel_ms,
                     "matches_golden": bool(sol.best_E == g["E"] and
                                            int("".join(map(str, sol.best_X)), 2) == g["state"])})
    wall_p = sum(r["wall_s"] for r in prod)
    return {"workload": "QBF on RandomQUBO-style n=24 int32, Q = triu(U)+triu(U,1)^T, U ~ integers[-100,100], "
                        "seeds 0-2, B=20 A=4 (2^24 states each), end to end incl. cache build",
            "impl": "numpy QBF restatement (oracle/qbf_numpy.py)", "cores": 1, "host_cpu_count": os.cpu_count(),
            "wall_s": wall, "states_per_s": 3 * 2.0 ** 24 / wall, "instances": rows,
            "all_match_golden": all(r["matches_golden"] for r in rows),
            "product_cpu": {"impl": "QBF().solve(p, device='cpu'): libqbx_cpu.so, the GPU's lane walk built for "
                                    "the host, OpenMP over lanes; best of 3 solves per instance",
                            "threads": HOST_THREADS,
                            "wall_s": wall_p, "states_per_s": 3 * 2.0 ** 24 / wall_p, "instances": prod,
                            "all_match_golden": all(r["matches_golden"] for r in prod)}}


# ------------------------------------------------------------------ exhaustive solves
def _golden(key):
    gold = ROOT / "tests" / "golden" / "oracle_goldens.json"
    if not gold.exists():
        return None
    return json.loads(gold.read_text()).get(key)


def _parity(E, state, key) -> dict:
    parity = {"E": E, "state": state}
    g = _golden(key)
    if g is not None:
        parity["oracle_state"] = g["state"]
        parity["oracle_E"] = g["E"]
        parity["abs_err_vs_oracle"] = abs(float(g["E"]) - float(E))
        parity["matches_oracle"] = bool(g["state"] == state and
                                        (g["E"] == E if isinstance(E, int) else abs(g["E"] - E) < 1e-9 * max(1, abs(E))))
    return parity


def _time_solves(prep, steps, warmup, kw):
    from qubo_lib_amd import engine
    for _ in range(warmup):
        engine.solve_matrix(prep, **kw)
    kms, results = [], []
    # every libqbx call ends with hipStreamSynchronize on its stream, so the clock below
    # brackets completed device work (torch is not loaded: it carries its own HIP runtime)
    t0 = time.perf_counter()
    for _ in range(steps):
        r = engine.solve_matrix(prep, **kw)
        kms.append(r.info["kernel_ms"])
        results.append((r.E, r.state))
    wall = time.perf_counter() - t0
    assert len(set(results)) == 1, "answers differ between steps"
    return r, kms, wall


def float_leg(args, devinfo, n: int = 34, steps: int = 500, warmup: int = 50) -> dict:
    """The float configuration: n = 34 N(0,1) float32 instance, FP64-exact answer.  A solve
    takes half a millisecond, so the leg times 500 of them (a quarter second) after 50
    untimed ones: over 20 steps after 3 the GPU's clock is still rising (0.51 against 0.48
    ms per solve measured)."""
    from qubo_lib_amd import engine
    from qubo_lib_amd.problems.generators import symmetric_normal_qubo
    Q = symmetric_normal_qubo(n, args.seed)
    r, kms, wall = _time_solves(engine.prepare(Q), steps, warmup, {})
    value = 2.0 ** n * steps / (float(np.sum(kms)) * 1e-3)
    parity = _parity(r.E, r.state, f"sym_normal_n{n}_s{args.seed}")
    return {"workload": f"symmetric N(0,1) float32 QUBO n={n}, seed {args.seed}: filter walk over quantized "
                        "tables, FP64 exact re-walks", "steps": steps, "warmup": warmup,
            "states_per_s": value, "kernel_ms_mean": float(np.mean(kms)), "ms_per_step": wall * 1e3 / steps,
            "walk": r.info["walk"], "walk_bits": r.info["walk_bits"], "tiles": r.info["tiles"],
            "abs_err_vs_fp64": parity.get("abs_err_vs_oracle"),
            "argmin_agrees": parity.get("oracle_state") == r.state if "oracle_state" in parity else None,
            "parity": parity,
            "roofline": roofline(value, r.info["walk"], r.info["inner_bits"], devinfo,
                                 "float" if n == 34 and args.seed == 0 else None)}


def float_ties(args, devinfo) -> dict:
    """Float inputs with ties in real arithmetic that FP64 keeps only to rounding (MaxCut
    G(n, 1/2) QUBOs scaled by 0.1 and by 1/3, off any power-of-two lattice; a float32 TSP
    of 6 cities, n = 36): device time per solve of the packed and the int32 float filter,
    best of --steps, and whether they agree (the default filter for a lone problem of
    n >= 28 is the packed one)."""
    from qubo_lib_amd import engine
    from qubo_lib_amd.problems.generators import maxcut_gnp
    from qubo_lib_amd.problems.tsp import TSP
    cases = []
    for n in (34, 36, 40):
        Q = maxcut_gnp(n, 0.5, seed=args.seed + 1).astype(np.float64)
        cases += [(f"maxcut G({n},1/2) x 0.1", Q * 0.1), (f"maxcut G({n},1/2) / 3", Q / 3)]
    cases.append(("TSP 6 cities float32 (seed 42)", TSP.init_random_instance(seed=42, n_cities=6).build_Q_matrix()))
    rows = []
    for name, Q in cases:
        p = engine.prepare(Q)
        row = {"instance": name, "n": p.n}
        answers = set()
        for f in ("packed", "int32"):
            r, kms, _ = _time_solves(p, max(args.steps, 1), max(args.warmup, 1), {"float_filter": f})
            row[f"{f}_kernel_ms"] = float(np.min(kms))
            answers.add((r.E, r.state))
        row["same_answer"] = len(answers) == 1
        row["E"] = r.E
        log(json.dumps(row))
        rows.append(row)
    return {"metric": "float filter device time per solve on tie-heavy inputs", "unit": "ms",
            "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "data": "synthetic", "dtype": "f64", "config": {"workload": "float-ties", "rows": rows}}


def exhaustive(args, ctx, devinfo):
    from qubo_lib_amd import engine
    from qubo_lib_amd.problems.generators import maxcut_gnp, symmetric_int_qubo, symmetric_normal_qubo
    if args.workload == "float":
        n = args.float_n
        Q = symmetric_normal_qubo(n, args.seed)
        workload = f"symmetric N(0,1) float32 QUBO n={n}, seed {args.seed}: filter walk over quantized tables, FP64 exact re-walk"
        dtype, key, gkey = "f32", ("float" if n == 34 else None), f"sym_normal_n{n}_s{args.seed}"
        metric = f"states/s at n={n} float32 (FP64-exact argmin)"
    elif args.workload == "maxcut":
        n = 40
        Q = maxcut_gnp(n, 0.5, seed=args.seed)
        workload = (f"MaxCut G(n=40, p=0.5) unit weights (seed {args.seed}), 2^40 states, exact: energy bound "
                    f"<= 15359, packed pair walk (two 16-bit states per 32-bit op)")
        dtype, key, gkey = "int32", "maxcut", f"maxcut_n40_s{args.seed}"
        metric = "states/s at n=40 MaxCut G(40, 0.5) (exact, packed pair walk)"
    else:
        n = 40
        Q = symmetric_int_qubo(n, args.seed, -127, 127, np.int16)
        workload = f"RandomQUBO n={n}, symmetric int16 U[-127,127] (seed {args.seed}), 2^{n} states, exact"
        dtype, key, gkey = "int32", "exhaustive", f"sym_int16_n40_s{args.seed}"
        metric = METRIC
    prep = engine.prepare(Q)
    kw = dict(walk_bits=args.walk_bits, inner_bits=args.inner_bits)
    r, kms, wall = _time_solves(prep, args.steps, args.warmup, kw)
    states = 2.0 ** n
    ms_per_step = wall * 1000.0 / args.steps
    kernel_ms = float(np.mean(kms))
    # value: Q already resident in HBM -- device time of each solve's kernels (HIP events on
    # the library's stream around prep + walk); the wall-clock rate with the per-solve
    # 3.2 KB host-to-device copy and 32 B result read-back is reported beside it
    value = states * args.steps / (float(np.sum(kms)) * 1e-3)
    value_host = states * args.steps / wall
    packed = r.info["walk"] in PACKED_WALKS
    E, state = r.E, r.state
    parity = _parity(E, state, gkey)
    if args.workload == "maxcut":
        # parity against the int32 walk on the same problem scaled out of the packed walks'
        # range (same argmin, energies times k)
        k = 4096   # couplings of 8192 per edge: too large for either packed walk
        ref = engine.solve_matrix(Q * k, **kw)
        parity.update(walk=r.info["walk"], int32_walk={"E": ref.E, "state": ref.state, "scale": k,
                                                       "walk": ref.info["walk"], "kernel_ms": ref.info["kernel_ms"]},
                      matches_int32_walk=bool(ref.E == E * k and ref.state == state))
    line = {
        "metric": metric,
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
        "roofline": roofline(value, r.info["walk"], r.info["inner_bits"], devinfo, key),
        "parity": parity,
        "host_inclusive": {"value": value_host, "unit": "states/s",
                           "note": "wall clock around the K solve calls: staging Q from host memory, "
                                   "launches, result read-back and synchronisation included"},
    }
    line["roofline"]["note"] = "kernel time = HIP events around prep+walk on the library stream"
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


# --------------------------------------------------------------------------- batches
def batch_leg(args, ctx, steps: int, warmup: int, plain: bool = True, generated: bool = True) -> dict:
    """4096-problem uneven mix: one persistent launch (graph replay) vs one launch per problem."""
    from qubo_lib_amd import engine
    from qubo_lib_amd.problems.generators import batch_mix
    probs, kinds = batch_mix(args.batch_count, seed=args.seed)
    Qs = [p.build_Q_matrix() for p in probs]
    preps = [engine.prepare(q) for q in Qs]
    mats = [p.matrix for p in preps]
    packed = ctx.pack(mats, np.int32)
    # batched form: the batch shape captured once in a HIP graph, replayed per step; each
    # problem's latency in this form is the wall time of the replay that solves it
    plan = ctx.batch_plan([m.shape[0] for m in mats], kind="i32")
    for _ in range(warmup):
        plan.run(packed=packed)
    walls = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        res = plan.run(packed=packed)
        walls.append(time.perf_counter() - t1)
    wall_b = (time.perf_counter() - t0) / steps
    graph_ms = float(res[0].kernel_ms)
    plan.close()
    # the same launches issued eagerly (no graph)
    for _ in range(warmup):
        ctx.batch(mats, kind="i32", packed=packed)
    t0 = time.perf_counter()
    for _ in range(steps):
        res_e = ctx.batch(mats, kind="i32", packed=packed)
    wall_e = (time.perf_counter() - t0) / steps
    # plain form: one C-ABI solve (launch + sync) per problem, same kernels and GPU
    lat, out_p = [], []
    t0 = time.perf_counter()
    for m in (mats if plain else []):
        t1 = time.perf_counter()
        out_p.append(ctx.solve_i32(m, walk_bits=None))
        lat.append(time.perf_counter() - t1)
    wall_p = time.perf_counter() - t0
    if not plain:          # profiling runs: batched launches only
        out_p, lat, wall_p = res, [float("nan")], float("nan")
    same = all((a.E_int, a.state) == (b.E_int, b.state) == (e.E_int, e.state)
               for a, b, e in zip(res, out_p, res_e))
    states = float(sum(2.0 ** q.shape[0] for q in Qs))
    out = {
        "workload": f"batch_mix({len(Qs)}, seed={args.seed}): MaxCut G(n,0.5) n~U{{16..32}} / int TSP 4-5 cities",
        "count": len(Qs), "states": states, "steps": steps, "warmup": warmup,
        "form": "one HIP graph replay per step (copy in, prep, persistent walks, copy out)",
        "solves_per_s": len(Qs) / wall_b, "ms_per_step": wall_b * 1e3, "graph_ms": graph_ms,
        "p50_ms": float(np.percentile(walls, 50) * 1e3), "p99_ms": float(np.percentile(walls, 99) * 1e3),
        "eager": {"solves_per_s": len(Qs) / wall_e, "ms_per_step": wall_e * 1e3},
        "plain": {"solves_per_s": len(Qs) / wall_p, "wall_s": wall_p,
                  "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
                  "form": "one qbx_solve_i32 (launch + sync) per problem, same kernels, tuned single tiling"},
        "speedup_vs_plain": wall_p / wall_b,
        "identical_answers": same,
        "states_per_s": states / wall_b,
        "tuning": str(ctx.tuning_file.name) if ctx.tuning_file else "built-in rule",
    }
    if generated:
        out["random_qubo_drawn_on_device"] = device_drawn_random_batch(args, ctx)
        out["tsp_built_on_device"] = device_built_tsp_batch(args, ctx)
    return out


def batch(args, ctx, devinfo):
    b = batch_leg(args, ctx, args.steps, args.warmup, plain=args.batch_plain)
    line = {"metric": "batched solves/s (4096-instance uneven mix)", "value": b["solves_per_s"], "unit": "solves/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": b["ms_per_step"],
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic", "config": {"workload": b["workload"], "states": b["states"], "form": b["form"],
                                            "tuning": b["tuning"]}}
    line.update({k: v for k, v in b.items() if k not in line})
    return line


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


# --------------------------------------------------------------------------- serving
def _serve_at(preps, arrivals, rate_scale, answers, warm=64):
    """Submit every request at its arrival time (scaled) to a fresh server; returns the
    rate row.  answers[j] collects request j's (E, state)."""
    from qubo_lib_amd.serve import Server
    arr = arrivals * rate_scale
    with Server(max_n=40) as srv:
        # warm-up: kernels loaded, slots touched
        for p in preps[:warm]:
            srv.submit(p)
        w = 0
        while w < warm:
            w += len(srv.poll(wait_ms=20))
        st0 = srv.stats()
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
    return {"solves_per_s": len(ids) / wall, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "wall_s": wall, "steps": int(st["steps"] - st0["steps"]),
            "device_busy_ms": st["busy_ms"] - st0["busy_ms"]}


def serve_leg(args, rates, count: int, burst: bool = True, sequential: bool = True) -> dict:
    """Poisson request stream (n in 20..36, P(n) ~ 2^-(n-20); half symmetric int16
    RandomQUBO in [-127,127], half MaxCut G(n,0.5)) served by the continuous-batching
    server at the given arrival rates, and by sequential plain solves at the first rate."""
    from qubo_lib_amd import engine
    from qubo_lib_amd.problems.generators import serving_requests
    from qubo_lib_amd.serve import Server
    base_rate = 20000.0
    reqs = serving_requests(count, base_rate, seed=args.seed)
    preps = [engine.prepare(q) for _, q, _ in reqs]
    arrivals = np.array([t for t, _, _ in reqs])
    rows = []
    answers = {}
    for rate in rates:
        row = {"rate_per_s": rate}
        row.update(_serve_at(preps, arrivals, base_rate / rate, answers))
        rows.append(row)
        log(f"serve rate={rate:.0f}/s -> {row}")
    out = {"workload": f"serving_requests({count}, seed={args.seed}): Poisson arrivals, n in 20..36, "
                       "P(n)~2^-(n-20), half int16 RandomQUBO [-127,127], half MaxCut G(n,0.5)",
           "rates": rows}
    if burst:
        # capacity: every request submitted at once -- from a buffer the client packed
        # beforehand (submit_packed), and from the prepared problems (submit_many, which packs)
        from qubo_lib_amd import _native
        assert all(p.kind == "i32" and p.integer for p in preps)
        t0 = time.perf_counter()
        packed = _native.Context.pack([p.matrix for p in preps], np.int32)
        pack_ms = (time.perf_counter() - t0) * 1e3
        for form in ("packed", "problems"):
            with Server(max_n=40) as srv:
                for p in preps[:64]:
                    srv.submit(p)
                w = 0
                while w < 64:
                    w += len(srv.poll(wait_ms=20))
                st0 = srv.stats()
                t0 = time.perf_counter()
                ids = list(srv.submit_packed(*packed)) if form == "packed" else srv.submit_many(preps)
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
            busy = st["busy_ms"] - st0["busy_ms"]
            row = {"solves_per_s": len(ids) / wall, "submit_us_per_request": t_sub / len(ids) * 1e6,
                   "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                   "wall_ms": wall * 1e3, "steps": int(st["steps"] - st0["steps"]),
                   "device_busy_ms": busy, "device_busy_frac": busy / (wall * 1e3)}
            if form == "packed":
                row["client_pack_ms"] = pack_ms
                out["burst"] = row
            else:
                out["burst_from_problems"] = row
            log(f"serve burst ({form}) -> {row}")
    if sequential:
        # sequential plain solves (one blocking launch each) on the same arrivals
        arr = arrivals * (base_rate / rates[0])
        t0 = time.perf_counter()
        lat = []
        for j, (a, p) in enumerate(zip(arr, preps)):
            while time.perf_counter() - t0 < a:
                pass
            r = engine.solve_matrix(p)
            lat.append((time.perf_counter() - t0 - a) * 1e3)
            answers.setdefault(j, set()).add((r.E, r.state))
        wall = time.perf_counter() - t0
        lat = np.array(lat)
        out["sequential"] = {"rate_per_s": rates[0], "solves_per_s": len(lat) / wall,
                             "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                             "wall_s": wall}
    out["identical_answers"] = all(len(v) == 1 for v in answers.values())
    return out


def serve(args, ctx, devinfo):
    s = serve_leg(args, args.serve_rates, args.serve_count)
    head = max(s["rates"], key=lambda r: r["solves_per_s"] if r["p99_ms"] < 5.0 else 0.0)
    line = {"metric": "serving solves/s (Poisson arrivals, n 20..36) with p50/p99 latency",
            "value": head["solves_per_s"], "unit": "solves/s",
            "n_gpus": 1, "steps": args.serve_count, "warmup": 64, "ms_per_step": 1000.0 / head["solves_per_s"],
            "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
            "dtype": "int32", "data": "synthetic",
            "config": {"workload": s["workload"], "headline_rate_per_s": head["rate_per_s"]},
            "p50_ms": head["p50_ms"], "p99_ms": head["p99_ms"]}
    line.update({k: v for k, v in s.items() if k not in line})
    return line


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
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["exhaustive", "float", "maxcut", "batch", "serve", "copy",
                                          "float-ties"],
                    default="exhaustive")
    ap.add_argument("--headline-only", action="store_true",
                    help="exhaustive workload: skip the batch / serve / cpu_c1 / float34 legs")
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
    elif args.workload == "float-ties":
        line = float_ties(args, devinfo)
    else:
        line = exhaustive(args, ctx, devinfo)
        if args.workload == "exhaustive" and not args.headline_only:
            # the metric's second clause and the other configurations, timed in this run
            log("batch: batch_mix(4096) one launch vs one launch per problem ...")
            line["batch"] = batch_leg(args, ctx, steps=max(args.steps, 5), warmup=max(args.warmup, 2))
            log("serve: 20000 Poisson requests at 20000/s ...")
            line["serve"] = serve_leg(args, [20000.0], 20000)
            log("float34: n=34 float32 ...")
            line["float34"] = float_leg(args, devinfo)
            log("cpu_c1: numpy QBF at n=24 int32, seeds 0-2 ...")
            line["cpu_c1"] = cpu_c1()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
