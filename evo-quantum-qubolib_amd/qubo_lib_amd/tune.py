"""Autotuner for the walk tiling: per n, kind and mode, the number of bits each lane walks.

A solve of size n is cut into 2^(n-L) lanes (the lane prefix bits, n-L of them) that each
walk 2^L states.  Small L means many short lanes: more parallelism, more per-lane set-up.
Large L means few long lanes: less set-up, but a lone problem may not fill the GPU.  The
best L depends on n and on how many problems share the GPU, so it is measured:

  mode "single": one problem alone; time = device time of one solve (HIP events).
  mode "batch":  a batch of same-size problems in one persistent launch, as in batches
                 and serving steps; time = device time of the launch / problems.

For every n the tuner times every valid L (or those within --window of the built-in rule),
keeps the fastest and
writes a JSON table that Context loads at start (qubo_lib_amd/tuning/<arch>.json, or
$QUBO_LIB_AMD_TUNING).  The answers never depend on L; only the time does.

    python -m qubo_lib_amd.tune --out qubo_lib_amd/tuning/gfx950.json

The reference has no such step: QBF sizes its split by memory (qbf.py:430-466) and QB_F1
uses a fixed B = 18 (qbf1.py:94-95).
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

from . import _native
from .problems.generators import symmetric_int_qubo, symmetric_normal_qubo

MAX_N = 56


def _matrix(kind: str, n: int, seed: int) -> np.ndarray:
    if kind == "i32":
        return symmetric_int_qubo(n, seed, -127, 127, np.int16).astype(np.int32)
    return symmetric_normal_qubo(n, seed).astype(np.float64)


def candidates(ctx: _native.Context, kind: str, mode: str, n: int, window: int) -> tuple[int, list[int]]:
    """(built-in L, valid L values within `window` of it)."""
    base = ctx.plan_info(kind, mode, n).walk_bits
    out = []
    for L in range(base - window, base + window + 1):
        if L < 1 or L > n:
            continue
        try:
            info = ctx.plan_info(kind, mode, n, walk_bits=L)
        except (ValueError, OverflowError):
            continue
        if info.walk_bits == L:
            out.append(L)
    return base, out


def time_single(ctx, kind, n, L, reps, seed=0) -> float:
    Q = _matrix(kind, n, seed)
    solve = ctx.solve_i32 if kind == "i32" else ctx.solve_f64
    solve(Q, walk_bits=L)
    return min(solve(Q, walk_bits=L).kernel_ms for _ in range(reps))


def time_batch(ctx, kind, n, L, reps, states, seed=0) -> float:
    count = int(max(16, min(4096, states // (1 << n))))
    Qs = [_matrix(kind, n, seed + i) for i in range(count)]
    dtype = np.int32 if kind == "i32" else np.float64
    packed = ctx.pack(Qs, dtype)
    ctx.batch(Qs, kind=kind, packed=packed, walk_bits=L)
    best = min(ctx.batch(Qs, kind=kind, packed=packed, walk_bits=L)[0].kernel_ms for _ in range(reps))
    return best / count


def tune(ctx: _native.Context | None = None, kinds=("i32", "f64"), modes=("single", "batch"),
         ns=None, window: int = 64, reps: int = 3, batch_states: int = 1 << 35, log=None) -> dict:
    """Measure every (kind, mode, n) and return the table (also installed on ctx)."""
    ctx = ctx or _native.get_context(0)
    log = log or (lambda *a: None)
    for kind in kinds:
        for mode in modes:
            ctx.set_walk_table(kind, mode, [])       # measure against the built-in rule
    result = {"walk_bits": {}, "measured_ms": {}, "builtin": {}}
    info = ctx.device_info()
    result["device"] = {"name": info["name"], "compute_units": info["compute_units"],
                        "clock_khz": info["clock_khz"]}
    for kind in kinds:
        for mode in modes:
            key = f"{kind}/{mode}"
            table = [-1] * (MAX_N + 1)
            measured = {}
            builtin = {}
            n_list = ns if ns is not None else (range(12, 41) if mode == "single" else range(12, 37))
            for n in n_list:
                base, cands = candidates(ctx, kind, mode, n, window)
                builtin[str(n)] = base
                times = {}
                for L in cands:
                    if mode == "single":
                        times[L] = time_single(ctx, kind, n, L, reps)
                    else:
                        times[L] = time_batch(ctx, kind, n, L, reps, batch_states)
                best = min(times, key=lambda L: (times[L], abs(L - base)))
                # keep the built-in rule unless the measured gain is above timing noise (2%)
                if times[best] > 0.98 * times.get(base, float("inf")):
                    best = base
                table[n] = best
                measured[str(n)] = {str(L): round(t, 6) for L, t in times.items()}
                log(f"{key} n={n}: built-in L={base} {times.get(base, float('nan')):.4f} ms, "
                    f"best L={best} {times[best]:.4f} ms")
            result["walk_bits"][key] = table
            result["measured_ms"][key] = measured
            result["builtin"][key] = builtin
            ctx.set_walk_table(kind, mode, table)
    return result


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", type=Path, default=None, help="JSON table to write (default: tuning/<arch>.json)")
    ap.add_argument("--kinds", nargs="+", default=["i32", "f64"], choices=["i32", "f64"])
    ap.add_argument("--modes", nargs="+", default=["single", "batch"], choices=["single", "batch"])
    ap.add_argument("--n", nargs="+", type=int, default=None)
    ap.add_argument("--window", type=int, default=64, help="largest |L - built-in L| tried")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args(argv)
    ctx = _native.Context(0, tuning=None)
    t0 = time.time()
    res = tune(ctx, args.kinds, args.modes, args.n, args.window, args.reps,
               log=lambda m: print(m, flush=True))
    res["seconds"] = round(time.time() - t0, 1)
    out = args.out or (_native.TUNING_DIR / f"{res['device']['name'].split(':')[0]}.json")
    out.parent.mkdir(parents=True, exist_ok=True)
    base = out if out.exists() else _native.TUNING_DIR / out.name
    if base.exists():
        # merge: keep entries of kinds/modes/n not measured in this run
        old = json.loads(base.read_text())
        for key, table in res["walk_bits"].items():
            prev = old.get("walk_bits", {}).get(key)
            if prev is not None:
                measured = {int(n) for n in res["measured_ms"][key]}
                res["walk_bits"][key] = [table[n] if n in measured else prev[n] for n in range(MAX_N + 1)]
        for sect in ("walk_bits", "measured_ms", "builtin"):
            merged = old.get(sect, {})
            for key, val in res[sect].items():
                if sect != "walk_bits" and isinstance(merged.get(key), dict):
                    merged[key].update({str(k): v for k, v in val.items()})
                else:
                    merged[key] = val
            res[sect] = merged
    out.write_text(json.dumps(res, indent=1, sort_keys=True))
    print(f"wrote {out} ({res['seconds']} s)", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
