"""Batched exhaustive solving: many problems of uneven N in one persistent launch.

`solve_batch(problems)` replaces a loop of `QBF().solve(p)` calls.  All problems of one
arithmetic path go to the GPU together: libqbx.so builds every problem's tables in one
launch and walks the 64-lane tiles of all problems from one persistent work queue, in
longest-tile-first order, so large and small problems share the GPU without per-problem
launches or host round trips.  Results (energies, argmins, tie rule) are identical to
solving each problem alone.
"""
from __future__ import annotations

import json

import numpy as np

from . import _native, engine
from .abc import ProblemABC
from .solution import Solution
from .state_vectors import StateVectors


def _device_family(problems):
    """'random_qubo' for seeded RandomQUBO instances sharing bounds and a dtype the device
    generator draws, 'tsp' for seeded random TSP instances (1..7 cities) sharing bounds and
    a float dtype, else None."""
    from ._native import GEN_DTYPES
    from .problems.random_qubo import RandomQUBO
    from .problems.tsp import TSP
    if not problems:
        return None
    first = problems[0]
    key = lambda p: (getattr(p, "low", None), getattr(p, "high", None), np.dtype(p.dtype))
    if all(isinstance(p, TSP) and getattr(p, "seed", None) is not None and getattr(p, "low", None) is not None
           and 1 <= int(p.n_cities) <= 7 and np.dtype(p.dtype).name in ("float32", "float64")
           and key(p) == key(first) for p in problems):
        return "tsp"
    for p in problems:
        if not isinstance(p, RandomQUBO) or getattr(p, "seed", None) is None:
            return None
        if np.dtype(p.dtype).name not in GEN_DTYPES:
            return None
        if key(p) != key(first):
            return None
    if np.issubdtype(np.dtype(first.dtype), np.integer):
        # integer draws are solved on the int32 path: its energy bound must hold for any draw
        n = max(int(p.N) for p in problems)
        m = max(abs(int(first.low)), abs(int(first.high) - 1))
        if m * (n + n * (n - 1)) > engine.INT32_BOUND:
            return None
    return "random_qubo"


def _drawable_on_device(problems) -> bool:
    """Whether the device can build these problems' matrices itself (_device_family)."""
    return _device_family(problems) is not None


def solve_batch(problems, device: str = "cuda", B=None, walk_bits=None, inner_bits=None,
                draw_on_device=None) -> list:
    """Solve every problem exactly; returns one Solution per problem, in order.  On a
    'cuda' device all problems of one arithmetic path go in one persistent launch; on
    'cpu' they are solved one after another by the host backend.

    draw_on_device: seeded RandomQUBO and random TSP batches can have their matrices built
    on the GPU (bit for bit the reference's, random_qubo.py:19-33, tsp.py:50-75, 272-321)
    instead of on the host; None selects it for such batches of 64 or more problems."""
    dev = _native.parse_device(device)
    family = _device_family(problems)
    if dev == "cpu":
        if draw_on_device:
            raise ValueError("draw_on_device needs a 'cuda' device")
        draw_on_device = False
    if draw_on_device and family is None:
        raise ValueError("draw_on_device needs seeded RandomQUBO or random TSP problems with one (low, high, dtype)")
    if draw_on_device is None:
        draw_on_device = family is not None and len(problems) >= 64
    if draw_on_device:
        ctx = _native.get_context(dev)
        p0 = problems[0]
        seeds = [int(p.seed) for p in problems]
        kw = dict(tie_B=B, walk_bits=walk_bits, inner_bits=inner_bits)
        if family == "tsp":
            ncs = [int(p.n_cities) for p in problems]
            res = ctx.random_tsp_batch(seeds, ncs, p0.low, p0.high, p0.dtype, **kw)
            mats = ctx.random_tsp_read(seeds, ncs, p0.low, p0.high, p0.dtype)   # for the Ising offsets
            integer = False
        else:
            ns = [int(p.N) for p in problems]
            res = ctx.random_qubo_batch(seeds, ns, p0.low, p0.high, p0.dtype, **kw)
            mats = ctx.random_qubo_read(seeds, ns, p0.low, p0.high, p0.dtype)   # for the Ising offsets
            integer = np.issubdtype(np.dtype(p0.dtype), np.integer)
        raws = [engine.raw_from_native(r, integer=integer) for r in res]
        mats = [m.astype(p0.dtype) for m in mats]    # the dtype build_Q_matrix returns
    else:
        mats = [np.asarray(p.build_Q_matrix(use_gpu=False)) for p in problems]
        raws = engine.solve_matrices(mats, device=dev, tie_B=B, walk_bits=walk_bits, inner_bits=inner_bits)
    params = json.dumps(dict(device=device, B=B, walk_bits=walk_bits, inner_bits=inner_bits))
    out = []
    for p, Q, r in zip(problems, mats, raws):
        n = Q.shape[0]
        x = engine.state_bits(r.state, n)
        sol = Solution(n, 2 ** (n - r.tie_B), 1, p.to_json(), params)
        sol.state_vectors = StateVectors(wide_vector_matrix=x.reshape(1, n))
        sol.energies = np.array([np.float64(r.E)])
        sol.metrics = [{}]
        sol.solver_class = "QBF"
        sol.best_E = np.float64(r.E)
        sol.best_E_exact = r.E
        sol.best_X = x
        sol.ising_offset = ProblemABC.get_Q_Ising_offset(np, Q)
        if hasattr(p, "prefix_E"):
            sol.prefix_E = p.prefix_E
        sol.time_to_solve_ms = r.info["kernel_ms"]
        sol.kernel_ms = r.info["kernel_ms"]
        sol.run_info = dict(r.info, tie_B=r.tie_B, state=r.state, rank=r.rank)
        out.append(sol)
    return out
