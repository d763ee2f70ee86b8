"""QBF: exhaustive QUBO solver, drop-in for qubo_lib.solvers.QBF (src/qubo_lib/solvers/qbf.py:20-70).

`QBF().solve(problem, **kwargs) -> Solution` accepts the reference's ProblemABC objects
(and this package's mirrors of them) and returns a Solution with the reference's fields
(qbf.py:328-360).  The search runs on a HIP GPU (gfx950) through libqbx.so, or on the
host CPU through libqbx_cpu.so; the hot loop of the reference (qbf.py:188-230) and its
caches (qbf.py:246-289) are replaced by the lane-prefix walk (the gfx950 kernels, or the
same walk compiled for the host).

kwargs (qbf.py:52-59, 429-441):
  device             'cpu' (default, as the reference: the GPU's lane walk built for the
                     host, OpenMP over lanes, libqbx_cpu.so) | 'cuda' | 'cuda:N' (HIP).
                     Both give the same answers.
  A, B               the reference's prefix/suffix split.  Here it only fixes the tie rule:
                     among equal energies the answer is the state the reference QBF would
                     return with that B (first prefix in Gray order, then lowest suffix).
                     Default B = min(N, 20), the reference CPU default.
  use_mem_GB         if given (and A, B are not), B is derived from it exactly as the
                     reference does (adjust_AB_to_fit_memory, qbf.py:455-466).
  correction_period  accepted for compatibility and not needed: the float path filters
                     in exact int32 arithmetic (no drift to correct) and computes every
                     energy it compares in FP64 from scratch.
  use_metrics        'time_arr' / 'energy_arr' / 'all' (qbf.py:342-351): per reference
                     prefix, in Gray visiting order, its least energy (energy_arr) and the
                     device (or host thread) time the walk spent on it (time_arr), both
                     recorded by the same walk that finds the answer.
  walk_bits, inner_bits  tuning overrides (bits walked per lane, unrolled block bits).
Obsolete kwargs use_gpu / gpu_id / devices raise ValueError as in the reference.
"""
from __future__ import annotations

import numpy as np

from .. import engine
from ..abc import ProblemABC, SolverABC
from ..solution import Solution
from ..state_vectors import StateVectors
from ..utils import PerformanceTimer

ALL_METRICS = {"time_arr", "energy_arr"}


def estimate_memory_consumption_bytes(A: int, B: int, sizeof: int) -> int:
    """qbf.py:448-452."""
    return (B * 2 ** B + A * 2 ** B) * sizeof


def adjust_AB_to_fit_memory(N: int, use_mem_GB: float, sizeof: int = 4):
    """qbf.py:454-466: the largest B <= N-2 whose tables fit use_mem_GB."""
    mem = use_mem_GB * 2 ** 30
    bestB = N - 2
    for B in range(N - 2, 0, -1):
        bestB = B
        if estimate_memory_consumption_bytes(N - B, B, sizeof) <= mem:
            break
    return N - bestB, bestB


def define_chunk_size(N: int, **kwargs):
    """(A, B) exactly as qbf.py:429-446 chooses them on the CPU path."""
    if "A" in kwargs:
        return kwargs["A"], N - kwargs["A"]
    if "B" in kwargs:
        return N - kwargs["B"], kwargs["B"]
    if "use_mem_GB" in kwargs:
        return adjust_AB_to_fit_memory(N, kwargs["use_mem_GB"], kwargs.get("size_of", 4))
    B = min(N, 20)
    return N - B, B


class QBF(SolverABC):
    solver_class = "QBF"

    def solve(self, problem: ProblemABC, **kwargs) -> Solution:
        self.prepare(problem, **kwargs)
        kw = dict(device=self.device_id, tie_B=self.B, walk_bits=kwargs.get("walk_bits"),
                  inner_bits=kwargs.get("inner_bits"))
        self.prefix_metrics = None
        if self.metrics_wanted(kwargs) and self.B >= 1:
            # one walk: the answer and the per-prefix records together
            res, minima, wave_ms = engine.solve_matrix_minima(self._prepared, **kw)
            self.prefix_metrics = (minima, wave_ms)
        else:
            res = engine.solve_matrix(self._prepared, **kw)
        self.timer.pick_time("Kernel run time")
        return self.pack_results_to_solution_instance(problem, res, **kwargs)

    @staticmethod
    def metrics_wanted(kwargs) -> set:
        use_metrics = kwargs.get("use_metrics", set())
        if use_metrics == "all":
            return set(ALL_METRICS)
        if isinstance(use_metrics, str):
            return {use_metrics}
        return set(use_metrics)

    def prepare(self, problem: ProblemABC, **kwargs):
        self.check_obsolete_args(**kwargs)
        self.kwargs = kwargs
        self.json_str = self.encode_kwargs(kwargs)
        self.parse_device_arg(**kwargs)
        self.timer = PerformanceTimer(self.use_gpu)
        self.N = problem.N
        self.Q = np.asarray(problem.build_Q_matrix(use_gpu=False))
        self.timer.pick_time("build_Q")
        if self.Q.shape != (self.N, self.N):
            raise ValueError(f"problem.N={self.N} but Q has shape {self.Q.shape}")
        kw = dict(kwargs)
        kw.setdefault("size_of", self.Q.dtype.itemsize)
        self.A, self.B = define_chunk_size(self.N, **kw)
        if not (0 <= self.B <= self.N):
            raise ValueError(f"B={self.B} outside [0, N={self.N}]")
        self.correction_period = kwargs.get("correction_period", 1024)
        self._prepared = engine.prepare(self.Q)
        self.timer.pick_time("prepare")

    def pack_results_to_solution_instance(self, problem, res: engine.RawResult, **kwargs) -> Solution:
        """Fields as qbf.py:328-360.  best_E is a float64 like the reference's (its
        prefix energy is float64); best_E_exact keeps the Python int for integer Q."""
        n = self.N
        x = engine.state_bits(res.state, n)
        sol = Solution(n, 2 ** self.A, 1, problem.to_json(), self.json_str)
        sol.state_vectors = StateVectors(wide_vector_matrix=x.reshape(1, n))
        E = np.float64(res.E)
        sol.energies = np.array([E])
        use_metrics = self.metrics_wanted(kwargs)
        metric_dict = {}
        if use_metrics:
            minima, wave_ms = self.prefix_metrics if self.prefix_metrics is not None else self.state_metrics()
            if "time_arr" in use_metrics:
                # the reference times each of its 2^A host-loop steps (qbf.py:362-365); here
                # all prefixes are walked at once, and each gets the time spent on its states
                metric_dict["time_arr"] = np.asarray(wave_ms, dtype=np.float64)
            if "energy_arr" in use_metrics:
                metric_dict["energy_arr"] = np.asarray(minima, dtype=np.float64)
        sol.metrics = [metric_dict]
        sol.solver_class = self.solver_class
        sol.best_E = E
        sol.best_E_exact = res.E
        sol.best_X = x
        sol.ising_offset = ProblemABC.get_Q_Ising_offset(np, self.Q)
        if hasattr(problem, "prefix_E"):
            sol.prefix_E = problem.prefix_E
        sol.time_to_solve_ms = self.timer.get_time_from_start()
        sol.timer = [(t, l) for t, l in self.timer.times if l != "epoch"]
        sol.kernel_ms = res.info["kernel_ms"]
        sol.run_info = dict(res.info, tie_B=res.tie_B, state=res.state, rank=res.rank)
        return sol

    STATE_METRICS_MAX_N = 20

    def state_metrics(self):
        """Metrics for B = 0, where every reference prefix is one state: the energies of
        the 2^N states in Gray visiting order (N <= 20), and no per-prefix time."""
        N = self.N
        if N > self.STATE_METRICS_MAX_N:
            raise ValueError(f"use_metrics with B=0 lists all 2^N energies: N={N} exceeds "
                             f"{self.STATE_METRICS_MAX_N}")
        i = np.arange(2 ** N, dtype=np.int64)
        g = i ^ (i >> 1)
        X = ((g[:, None] >> np.arange(N - 1, -1, -1, dtype=np.int64)) & 1).astype(np.float64)
        Qf = np.asarray(self.Q, dtype=np.float64)
        E = ((X @ Qf) * X).sum(axis=1)
        return E, np.zeros(2 ** N)

    def clear_mem(self):
        self.Q = None
        self._prepared = None
