"""Problem with fixed leading variables, mirroring qubo_lib's ReducedProblem
(src/qubo_lib/problems/reduced_problem.py:5-57).

Fixing x[:A] = prefix gives the constant prefix_E = p^T Q_AA p and the reduced matrix
Q' = Q_BB + diag(p (Q_AB + Q_BA^T)) (reduced_problem.py:31-46).  The same numpy
expressions are evaluated in the problem's dtype, so results (including any integer
wrap-around the reference would show) are identical.
"""
from __future__ import annotations

import numpy as np

from ..abc import ProblemABC


class ReducedProblem(ProblemABC):
    def __init__(self, problem, prefix):
        prefix = np.asarray(prefix)
        self.base_problem = problem
        self.A = prefix.shape[0]
        self.prefix = prefix
        self.prefix_E = None
        super().__init__(problem.N - self.A, problem.dtype)

    def build_Q_matrix(self, use_gpu: bool = False) -> np.ndarray:
        Q = np.asarray(self.base_problem.build_Q_matrix(use_gpu=False))
        A = self.A
        x = np.array(self.prefix, dtype=self.dtype)
        self.prefix_E = (x @ Q[:A, :A] @ x.T).item()
        dgnl = x @ (Q[:A, A:] + Q[A:, :A].T)
        result = Q[A:, A:] + np.diag(dgnl)
        assert result.dtype == self.dtype
        return result

    @staticmethod
    def init_random_instance(*args, **kwargs):
        raise NotImplementedError("It's a user-defined matrix, for random instance better use RandomQUBO class")
