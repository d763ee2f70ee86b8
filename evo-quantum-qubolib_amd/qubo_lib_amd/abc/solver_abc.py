"""Solver contract, mirroring qubo_lib's SolverABC (src/qubo_lib/abc/solver_abc.py:18-92).

Solvers take no constructor arguments (solver_abc.py:26) and are driven entirely by
keyword arguments of `solve(problem, **kwargs) -> Solution` (solver_abc.py:37-60).
`device` is 'cpu' (the default, as in the reference: the lane walk on the host CPU,
libqbx_cpu.so), 'cuda' or 'cuda:N' (a HIP device, libqbx.so) (solver_abc.py:62-75).
"""
from __future__ import annotations

import json
from abc import ABC, abstractmethod

import numpy as np

from .. import _native
from ..solution import Solution


class SolverABC(ABC):
    def __init__(self):
        self.use_gpu = None
        self.kwargs = None
        self.json_str = None
        self.timer = None
        self.N = None
        self.Q = None
        self.epochs_cnt = None
        self.samples_cnt = None
        self.device_name = None
        self.device_id = None

    @abstractmethod
    def solve(self, problem, **kwargs) -> Solution:
        """Solve `problem`; every option is a keyword argument."""

    @staticmethod
    def check_obsolete_args(**kwargs):
        """Same rejections as the reference (qbf.py:105-113)."""
        if "use_gpu" in kwargs:
            raise ValueError("use_gpu param is obsolete!")
        if "gpu_id" in kwargs:
            raise ValueError("gpu_id param is obsolete!")
        if "devices" in kwargs:
            raise ValueError("Use device argument")

    def parse_device_arg(self, default: str = "cpu", **kwargs):
        """'cpu' (the reference's default) -> the host CPU backend, 'cuda' -> HIP device 0,
        'cuda:N' -> device N (solver_abc.py:62-75); anything else raises ValueError."""
        device = kwargs.get("device", default)
        if not isinstance(device, str):
            raise ValueError(f"Unknown device name: {device}")
        if device == "cpu":
            self.use_gpu = False
            self.device_id = "cpu"
        elif device[:4] == "cuda":
            self.use_gpu = True
            self.device_id = _native.parse_device(device)
        else:
            raise ValueError(f"Unknown device name: {device}")
        self.device_name = device
        return self.device_id

    @staticmethod
    def encode_kwargs(kwargs) -> str:
        """JSON record of the solve kwargs (the reference uses jsonpickle, qbf.py:75-76)."""
        def enc(o):
            if isinstance(o, (set, frozenset)):
                return sorted(o)
            if isinstance(o, np.generic):
                return o.item()
            return repr(o)
        return json.dumps(kwargs, indent=4, default=enc, sort_keys=True)

    @staticmethod
    def gimme_lib():
        return np
