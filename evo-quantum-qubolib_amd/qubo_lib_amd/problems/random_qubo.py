"""Uniform random QUBO, mirroring qubo_lib's RandomQUBO (src/qubo_lib/problems/random_qubo.py:7-61).

Q is drawn from Generator(PCG64(seed)) exactly as the reference draws it
(random_qubo.py:19-33): integer dtypes take rng.integers(low, high, (N, N), dtype) with
`high` exclusive; float dtypes take low + rng.random((N, N), dtype) * (high - low).  The
matrix is not symmetrised.  Same seed, bounds and dtype give the same Q, bit for bit.
"""
from __future__ import annotations

import numpy as np
from numpy.random import PCG64, Generator

from ..abc import ProblemABC


class RandomQUBO(ProblemABC):
    def __init__(self, N: int, dtype=np.float32):
        super().__init__(N, dtype)
        self.low = -100
        self.high = 100

    def build_Q_matrix(self, use_gpu: bool = False) -> np.ndarray:
        shape = (self.N, self.N)
        rng = Generator(PCG64(self.seed))
        if np.issubdtype(self.dtype, np.integer):
            Q = rng.integers(low=self.low, high=self.high, size=shape, dtype=self.dtype)
        else:
            Q = self.low + rng.random(size=shape, dtype=self.dtype) * (self.high - self.low)
        assert Q.dtype == self.dtype
        return Q

    @staticmethod
    def init_random_instance(seed: int, N: int, low=-100, high=100, dtype=np.float32) -> "RandomQUBO":
        obj = RandomQUBO(N, dtype)
        obj.seed = seed
        obj.low = low
        obj.high = high
        return obj
