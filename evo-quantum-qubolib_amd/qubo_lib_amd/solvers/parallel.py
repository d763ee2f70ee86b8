"""Parallel: bit-fixing decomposition, drop-in for the bruteforce branch of
qubo_lib.solvers.Parallel (src/qubo_lib/solvers/parallel.py:20-158, 225-320).

The top k = log2(splits_cnt) variables are fixed to every value in turn, giving
splits_cnt ReducedProblems (parallel.py:145-157, reduced_problem.py:31-46).  Instead of one
spawned process per device with pickled tasks (scheduler.py:5-59), each device solves its
share of the reduced problems in ONE persistent batch launch, from a host thread per
device.  The merge takes the minimum of best_E + prefix_E (parallel.py:273-284); on exact
ties the lowest task id wins, deterministically, where the reference keeps whichever
result arrived first.

Sampler base solvers (MSSA) are not part of this build and are rejected.
"""
from __future__ import annotations

import threading
import time

import numpy as np

from .. import engine
from ..abc import SolverABC
from ..problems import ReducedProblem
from ..solution import Solution
from ..state_vectors import StateVectors
from .qbf import QBF
from .qbf1 import QB_F1

SAMPLER_SOLVERS: dict = {}
BRUTEFORCE_SOLVERS = dict(QBF=QBF, QB_F1=QB_F1)
SUPPORTED_SOLVERS = dict(**SAMPLER_SOLVERS, **BRUTEFORCE_SOLVERS)


class Parallel(SolverABC):
    def __init__(self, base_solver, splits_cnt: int, devices):
        self._base_solver = base_solver
        self._splits_cnt = splits_cnt
        self._devices = devices
        self._validate_config()

    def _validate_config(self):
        rev = {v: k for k, v in SUPPORTED_SOLVERS.items()}
        if self._base_solver not in rev:
            raise ValueError(f"Unsupported solver class: {self._base_solver}. "
                             f"Supported solvers: {', '.join(SUPPORTED_SOLVERS)}")
        self._solver_class_name = rev[self._base_solver]
        if not isinstance(self._splits_cnt, int) or self._splits_cnt <= 0:
            raise ValueError("splits_cnt must be a positive integer")
        if self._splits_cnt & (self._splits_cnt - 1) != 0:
            raise ValueError("For bruteforce solvers with Embarrassingly Parallel Enumeration / "
                             f"Bit-Fixing Decomposition, splits_cnt must be a power of 2, got {self._splits_cnt}")
        if not self._devices:
            raise ValueError("Devices list cannot be empty")
        has_cpu = any("cpu" in d for d in self._devices)
        has_gpu = any("cuda" in d for d in self._devices)
        if has_cpu and has_gpu:
            raise ValueError("Cannot mix CPU and GPU devices")
        if has_cpu:
            raise ValueError("qubo_lib_amd runs on HIP GPUs only; use devices=['cuda', ...]")
        if all(d == "cuda" for d in self._devices):
            self._devices = [f"cuda:{i}" for i in range(len(self._devices))]
        else:
            self._devices = list(self._devices)

    def solve(self, problem, timeout=None, **kwargs) -> Solution:
        SolverABC.check_obsolete_args(**kwargs)
        kw = dict(solver_type=self._solver_class_name, splits_cnt=self._splits_cnt,
                  devices=self._devices, timeout=timeout, **kwargs)
        json_str = self.encode_kwargs(kw)
        k = self._splits_cnt.bit_length() - 1
        if k > problem.N:
            raise ValueError(f"splits_cnt=2^{k} exceeds 2^N for N={problem.N}")
        reduced = []
        for task_id in range(self._splits_cnt):
            prefix = np.array([(task_id >> (k - 1 - j)) & 1 for j in range(k)], dtype=np.int64)
            reduced.append(ReducedProblem(problem, prefix))
        if self._solver_class_name == "QB_F1" and np.dtype(problem.dtype) != np.int16:
            raise TypeError("Q dtype only np.int16 supported yet")
        # tasks are dealt to devices round robin; each device solves its share in one launch
        ndev = len(self._devices)
        shares = [list(range(d, self._splits_cnt, ndev)) for d in range(ndev)]
        results: dict = {}
        errors: list = []

        def work(worker_id):
            dev = int(self._devices[worker_id][5:]) if len(self._devices[worker_id]) > 5 else 0
            ids = shares[worker_id]
            if not ids:
                return
            t0 = time.perf_counter()
            try:
                mats = [reduced[t].build_Q_matrix() for t in ids]
                n_sub = problem.N - k
                if n_sub == 0:
                    raws = [None] * len(ids)
                else:
                    Bsub = kwargs.get("B", None)
                    raws = engine.solve_matrices(mats, device=dev, tie_B=Bsub,
                                                 walk_bits=kwargs.get("walk_bits"),
                                                 inner_bits=kwargs.get("inner_bits"))
                ms = (time.perf_counter() - t0) * 1000.0
                for t, r in zip(ids, raws):
                    results[t] = (worker_id, r, ms / len(ids))
            except Exception as e:  # reported on the merged Solution, never silently dropped
                errors.append((worker_id, e))

        threads = [threading.Thread(target=work, args=(w,)) for w in range(ndev)]
        t_start = time.perf_counter()
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        wall_ms = (time.perf_counter() - t_start) * 1000.0
        if errors:
            raise errors[0][1]

        merged = Solution(N=problem.N, epochs_cnt=self._splits_cnt, samples_cnt=1,
                          problem_instance_json=problem.to_json(), solver_params_json=json_str)
        merged.ising_offset = None
        merged.solver_class = self._solver_class_name
        best = None
        for task_id in range(self._splits_cnt):
            worker_id, r, _ = results[task_id]
            pe = reduced[task_id].prefix_E
            e_sub = 0 if r is None else r.E
            total = e_sub + pe
            if best is None or total < best[0]:
                best = (total, task_id, r)
        total, task_id, r = best
        prefix = reduced[task_id].prefix.astype(np.int32)
        sub_x = np.zeros(0, dtype=np.int32) if r is None else engine.state_bits(r.state, problem.N - k)
        full = np.hstack([prefix, sub_x]).astype(np.int32)
        merged.best_E = total
        merged.energies = np.array([total])
        merged.best_X = full
        merged.state_vectors = StateVectors(wide_vector_matrix=full.reshape(1, -1))
        merged.metrics = [{}]
        worker_times: dict = {}
        for t, (w, _, ms) in results.items():
            worker_times[w] = worker_times.get(w, 0.0) + ms
        merged.time_to_solve_ms = max(worker_times.values()) if worker_times else 0.0
        merged.worker_logs = [dict(task_id=t, worker_id=w, device=self._devices[w], time_to_solve_ms=ms)
                              for t, (w, _, ms) in sorted(results.items())]
        if timeout is not None and wall_ms > timeout * 1000.0:
            merged.error = f"timeout: solve took {wall_ms:.1f} ms > {timeout} s"
        return merged
