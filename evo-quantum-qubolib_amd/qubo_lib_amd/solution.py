"""Result container, mirroring qubo_lib's Solution (src/qubo_lib/solution.py:3-85).

Same fields and printed form.  Serialization uses plain JSON (the reference uses
jsonpickle, solution.py:64-69, which is not available here); `init_from_json` restores
every field, including numpy arrays and the StateVectors.
"""
from __future__ import annotations

import json

import numpy as np

from .state_vectors import StateVectors


class Solution:
    def __init__(self, N: int, epochs_cnt: int, samples_cnt: int, problem_instance_json: str,
                 solver_params_json: str):
        self.N = N
        self.epochs_cnt = epochs_cnt
        self.samples_cnt = samples_cnt
        self.problem_instance_json = problem_instance_json
        self.solver_params_json = solver_params_json
        self.metrics = []
        self.state_vectors = None
        self.energies = None
        self.ising_offset = None
        self.solver_class = None
        self.best_E = None
        self.best_X = None
        self.error = ""
        self.time_to_solve_ms = None

    # -- serialization ------------------------------------------------------------------
    def to_json(self) -> str:
        return json.dumps(_encode(self.__dict__), indent=4)

    @staticmethod
    def init_from_json(json_str: str) -> "Solution":
        d = _decode(json.loads(json_str))
        obj = Solution.__new__(Solution)
        obj.__dict__.update(d)
        return obj

    def __str__(self):
        if self.error:
            return f"Error: {self.error}"
        best_x_str = "".join(map(str, self.best_X)) if self.best_X is not None else None
        return (
            "Solution(\n"
            f"  N={self.N},\n"
            f"  best_E={self.best_E:.3f},\n"
            f"  time_to_solve_ms={self.time_to_solve_ms:.1f},\n"
            f"  samples_cnt={self.samples_cnt},\n"
            f"  solver_class='{self.solver_class}',\n"
            f"  best_X='{best_x_str}'\n"
            ")"
        )


def _encode(o):
    if isinstance(o, np.ndarray):
        return {"__ndarray__": o.tolist(), "dtype": str(o.dtype), "shape": list(o.shape)}
    if isinstance(o, StateVectors):
        return {"__state_vectors__": _encode(o.bitwise_vectors), "N": o.N, "element_bits": o.element_bits}
    if isinstance(o, np.generic):
        return {"__npscalar__": o.item(), "dtype": str(o.dtype)}
    if isinstance(o, dict):
        return {str(k): _encode(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_encode(v) for v in o]
    return o


def _decode(o):
    if isinstance(o, dict):
        if "__ndarray__" in o:
            return np.array(o["__ndarray__"], dtype=o["dtype"]).reshape(o["shape"])
        if "__state_vectors__" in o:
            sv = StateVectors.__new__(StateVectors)
            sv.bitwise_vectors = _decode(o["__state_vectors__"])
            sv.N = o["N"]
            sv.element_bits = o["element_bits"]
            return sv
        if "__npscalar__" in o:
            return np.dtype(o["dtype"]).type(o["__npscalar__"])
        return {k: _decode(v) for k, v in o.items()}
    if isinstance(o, list):
        return [_decode(v) for v in o]
    return o
