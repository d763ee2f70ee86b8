"""QB_F1: int16-only exhaustive solver, drop-in for qubo_lib.solvers.QB_F1
(src/qubo_lib/solvers/qbf1.py:21-318).

Same contract as the reference: GPU only (a non-'cuda' device raises ValueError,
qbf1.py:308-318), int16 Q only (TypeError, qbf1.py:99-100), solver_class 'QB_F1',
epochs_cnt 1, integer best_E (qbf1.py:253-289).  It runs the same exact kernels as QBF.

Differences, by design: energies are exact (int32 accumulation) where the reference
accumulates in int16 and can wrap (qbf1.py:44, 158); any N in [1, 56] is accepted (the
reference's table construction needs 28 <= N <= 48); among equal energies the state
returned is the one QBF returns with B = 18, the reference QB_F1's suffix width
(qbf1.py:94).  The reference QB_F1 orders ties by its thread layout instead
(qbf1.py:191-247, 254), which has no meaning on a different launch geometry.
"""
from __future__ import annotations

import numpy as np

from .. import engine
from ..abc import ProblemABC
from ..solution import Solution
from ..state_vectors import StateVectors
from ..utils import PerformanceTimer
from .qbf import QBF


class QB_F1(QBF):
    solver_class = "QB_F1"
    SUFFIX_BITS = 18

    def solve(self, problem: ProblemABC, **kwargs) -> Solution:
        self.check_obsolete_args(**kwargs)
        self.json_str = self.encode_kwargs(kwargs)
        self.timer = PerformanceTimer(True)
        device = kwargs.get("device", "cuda")
        if not isinstance(device, str) or device[:4] != "cuda":
            raise ValueError(f"QB_F1 solver is GPU only: {device}")
        self.parse_device_arg(default="cuda", **kwargs)
        self.N = problem.N
        self.Q = np.asarray(problem.build_Q_matrix(use_gpu=False))
        if self.Q.dtype != np.int16:
            raise TypeError("Q dtype only np.int16 supported yet")
        self.timer.pick_time("build_Q")
        self.B = min(self.SUFFIX_BITS, self.N)
        self.A = self.N - self.B
        res = engine.solve_matrix(self.Q, device=self.device_id, tie_B=self.B,
                                  walk_bits=kwargs.get("walk_bits"), inner_bits=kwargs.get("inner_bits"))
        self.timer.pick_time("Kernel run time")
        n = self.N
        x = engine.state_bits(res.state, n)
        sol = Solution(n, 1, 1, problem.to_json(), self.json_str)
        sol.state_vectors = StateVectors(wide_vector_matrix=x.reshape(1, n))
        sol.energies = np.array([res.E])
        sol.metrics = [{}]
        sol.solver_class = self.solver_class
        sol.best_E = int(res.E)
        sol.best_E_exact = res.E
        sol.best_X = x
        if hasattr(problem, "prefix_E"):
            sol.prefix_E = problem.prefix_E
        sol.ising_offset = ProblemABC.get_Q_Ising_offset(np, self.Q)
        sol.timer = [(t, l) for t, l in self.timer.times if l != "epoch"]
        sol.time_to_solve_ms = self.timer.get_time_from_start()
        sol.kernel_ms = res.info["kernel_ms"]
        sol.run_info = dict(res.info, tie_B=res.tie_B, state=res.state, rank=res.rank)
        return sol
