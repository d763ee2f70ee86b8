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
  use_metrics        'time_arr' / 'energy_arr' / 'all' (qbf.py:342-351); see
                     pack_results for what each holds.
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
        res = engine.solve_matrix(self._prepared, device=self.device_id, tie_B=self.B,
                                  walk_bits=kwargs.get("walk_bits"), inner_bits=kwargs.get("inner_bits"))
        self.timer.pick_time("Kernel run time")
        return self.pack_results_to_solution_instance(problem, res, **kwargs)

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
        use_metrics = kwargs.get("use_metrics", set())
        if use_metrics == "all":
            use_metrics = ALL_METRICS
        metric_dict = {}
        if "time_arr" in use_metrics:
            # the reference times each of its 2^A host-loop steps; a single kernel walks
            # all prefixes at once, so its one device interval is reported
            metric_dict["time_arr"] = np.array([res.info["kernel_ms"]])
        if "energy_arr" in use_metrics:
            metric_dict["energy_arr"] = self.prefix_minima(problem)
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

    def prefix_minima(self, problem) -> np.ndarray:
        """energy_arr: for each of the 2^A reference prefixes, in Gray visiting order, the
        minimum energy over its 2^B suffixes (qbf.py:409-410).  Computed exactly on the
        GPU as 2^A reduced problems of size B in one batch."""
        from ..problems import ReducedProblem, Just_Q
        A, B = self.A, self.B
        if A > 16:
            raise ValueError(f"energy_arr needs 2^A problems; A={A} is too large (max 16)")
        base = Just_Q(self.Q)
        out = np.empty(2 ** A)
        subs, offs = [], []
        for i in range(2 ** A):
            g = i ^ (i >> 1)
            prefix = np.array([(g >> (A - 1 - j)) & 1 for j in range(A)], dtype=np.int64)
            rp = ReducedProblem(base, prefix)
            q = rp.build_Q_matrix()
            subs.append(q)
            offs.append(rp.prefix_E)
        if B == 0:
            return np.array(offs, dtype=np.float64)
        res = engine.solve_matrices(subs, device=self.device_id)
        for i, (r, pe) in enumerate(zip(res, offs)):
            out[i] = float(r.E) + float(pe)
        return out

    def clear_mem(self):
        self.Q = None
        self._prepared = None
