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

from . import engine
from .abc import ProblemABC
from .solution import Solution
from .state_vectors import StateVectors


def solve_batch(problems, device: str = "cuda", B=None, walk_bits=None, inner_bits=None) -> list:
    """Solve every problem exactly; returns one Solution per problem, in order."""
    dev = 0
    if isinstance(device, str) and device[:4] == "cuda":
        dev = int(device[5:]) if len(device) > 5 else 0
    else:
        raise ValueError(f"solve_batch runs on the GPU only: {device}")
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
