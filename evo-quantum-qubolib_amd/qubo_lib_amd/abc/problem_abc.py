"""Problem contract, mirroring qubo_lib's ProblemABC (src/qubo_lib/abc/problem_abc.py:6-89).

A problem knows its size N and dtype and builds its dense Q lazily.  Solvers only need
`N`, `dtype`, `build_Q_matrix(use_gpu)`, `to_json()` and, for reduced problems, the
`prefix_E` attribute that build_Q_matrix sets.  Objects of the reference's own problem
classes satisfy the same duck-typed contract and are accepted by every solver here.
"""
from __future__ import annotations

import json
from abc import ABC, abstractmethod
from typing import Optional

import numpy as np


class ProblemABC(ABC):
    def __init__(self, N: int, dtype=np.float32):
        self.N = N
        self.dtype = dtype
        self.seed = None
        self.ground_E = None
        self.ground_state = None
        self.ground_utility = None

    @abstractmethod
    def build_Q_matrix(self, use_gpu: bool = False) -> np.ndarray:
        """Dense N x N Q as a numpy array.  `use_gpu` is accepted for interface
        compatibility; device upload happens inside the solver (problem_abc.py:16-30)."""

    def to_json(self) -> str:
        return json.dumps(self._json_fields(), default=_json_default)

    def _json_fields(self) -> dict:
        d = {"py/object": f"{type(self).__module__}.{type(self).__qualname__}"}
        for k, v in self.__dict__.items():
            d[k] = v
        return d

    @staticmethod
    def init_from_json(json_str: str) -> "ProblemABC":
        raise NotImplementedError

    @staticmethod
    def init_random_instance(seed: int, N: int, dtype=np.float32) -> "ProblemABC":
        raise NotImplementedError

    def calculate_utility_metric(self, states) -> Optional[np.ndarray]:
        return None

    def calculate_ground(self):
        return None

    @staticmethod
    def get_Q_Ising_offset(lib, Q):
        """1/4 * sum(Q) + 1/4 * sum(diag(Q))  (problem_abc.py:86-89)."""
        return 1 / 4 * lib.sum(Q) + 1 / 4 * lib.sum(lib.diag(Q))


def _json_default(o):
    if isinstance(o, np.ndarray):
        return {"py/numpy.ndarray": {"dtype": str(o.dtype), "values": o.tolist()}}
    if isinstance(o, np.generic):
        return o.item()
    if isinstance(o, type):
        return {"py/type": f"{o.__module__}.{o.__qualname__}"}
    if isinstance(o, ProblemABC):
        return o._json_fields()
    return repr(o)
