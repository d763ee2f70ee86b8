"""User-given Q, mirroring qubo_lib's Just_Q (src/qubo_lib/problems/just_Q.py:5-38)."""
from __future__ import annotations

import numpy as np

from ..abc import ProblemABC


class Just_Q(ProblemABC):
    def __init__(self, Q):
        Q = np.asarray(Q)
        if Q.ndim != 2 or Q.shape[0] != Q.shape[1]:
            raise ValueError(f"Q must be a square matrix, got shape {Q.shape}")
        self.Q_matrix = Q
        super().__init__(Q.shape[0], Q.dtype)

    def build_Q_matrix(self, use_gpu: bool = False) -> np.ndarray:
        assert self.Q_matrix.dtype == self.dtype
        return self.Q_matrix

    @staticmethod
    def init_random_instance(*args, **kwargs):
        raise NotImplementedError("It's a user-defined matrix, for random instance better use RandomQUBO class")
