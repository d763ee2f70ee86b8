"""Parallel: bit-fixing decomposition, drop-in for the bruteforce branch of
qubo_lib.solvers.Parallel (src/qubo_lib/solvers/parallel.py:20-158, 225-320).

The top k = log2(splits_cnt) variables are fixed to every value in turn, giving
splits_cnt reduced problems (parallel.py:145-157; the algebra of reduced_problem.py:31-46).  Instead of one
spawned process per device with pickled tasks (scheduler.py:5-59), each GPU solves its
share of the reduced problems in ONE persistent batch launch, from a host thread per
device; 'cpu' devices solve their shares one after another on the host backend (each
solve already spreads over every core).

Merge.  The reference keeps the first minimum of best_E + prefix_E to arrive
(parallel.py:273-284), so its argmin among equal energies depends on process timing.  Here
the merge is the base solver's own tie rule over the whole problem: every task returns the
minimum of (E, full rank) over its states, and the tasks merge by (E, full rank), so
Parallel(QBF, 2^k).solve(p).best_X == QBF().solve(p).best_X whenever the energies are
exact.  For that, a task's sub-solve must order its states by their ranks in the FULL
problem, rank(s) = (invgray(s >> B) << B) | (s & (2^B - 1)) (qbf.py:212-230, 395-412):
  * k >= N - B: the fixed bits cover the reference prefix, the free bits are suffix bits,
    ranked in binary order: the sub-solve uses tie width B' = N - k (its whole size);
  * k < N - B: the free bits' top N - B - k are prefix bits, whose Gray ranks are
    complemented when the fixed bits have odd parity (invgray of the concatenation).
    Complementing the Gray rank of a prefix q is flipping q's top bit
    (gray(2^m - 1) = 2^(m-1)), so such a task solves the reduced matrix with its first
    variable replaced by its complement (x = 1 - y: an exact change of Q plus a constant)
    under the same tie width B; even-parity tasks solve the reduced matrix as it is.
Reduced matrices are built in int64 / float64 (exact for integer and float32 inputs)
rather than in the problem's dtype, so energies are not rounded to float32 on the way.

Sampler base solvers (MSSA) are not part of this build and are rejected.
"""
from __future__ import annotations

import threading
import time

import numpy as np

from .. import _native, engine
from ..abc import SolverABC
from ..solution import Solution
from ..state_vectors import StateVectors
from .qbf import QBF, define_chunk_size
from .qbf1 import QB_F1

SAMPLER_SOLVERS: dict = {}
BRUTEFORCE_SOLVERS = dict(QBF=QBF, QB_F1=QB_F1)
SUPPORTED_SOLVERS = dict(**SAMPLER_SOLVERS, **BRUTEFORCE_SOLVERS)


def _invgray(v: int) -> int:
    r = 0
    while v:
        r ^= v
        v >>= 1
    return r


def full_rank(state: int, B: int) -> int:
    """rank(s) of the reference tie rule (qbf.py:212-230, 395-412)."""
    return (_invgray(state >> B) << B) | (state & ((1 << B) - 1))


def reduced_matrix(Q: np.ndarray, prefix: np.ndarray):
    """(R, prefix_E) of fixing x[:k] = prefix: R = Q_BB + diag(p (Q_AB + Q_BA^T)),
    prefix_E = p^T Q_AA p (reduced_problem.py:31-46), in int64 or float64."""
    wide = np.int64 if np.issubdtype(Q.dtype, np.integer) else np.float64
    Q = Q.astype(wide)
    k = prefix.shape[0]
    x = prefix.astype(wide)
    pe = x @ Q[:k, :k] @ x
    R = Q[k:, k:] + np.diag(x @ (Q[:k, k:] + Q[k:, :k].T))
    return R, (int(pe) if wide is np.int64 else float(pe))


def complement_first(R: np.ndarray):
    """(R', c) with E_R(x) = E_R'(y) + c for y = x with its first bit complemented
    (x_0 = 1 - y_0): R'_00 = -R_00, R'_0j = -R_0j, R'_j0 = -R_j0, R'_jj += R_0j + R_j0."""
    Rp = R.copy()
    c = R[0, 0]
    side = R[0, 1:] + R[1:, 0]
    Rp[0, :] = -R[0, :]
    Rp[1:, 0] = -R[1:, 0]
    Rp[np.arange(1, R.shape[0]), np.arange(1, R.shape[0])] += side
    return Rp, (int(c) if np.issubdtype(R.dtype, np.integer) else float(c))


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
        if has_cpu and self._base_solver is QB_F1:
            raise ValueError("QB_F1 solver is GPU only: cpu")
        for d in self._devices:
            _native.parse_device(d)
        if all(d == "cuda" for d in self._devices):
            self._devices = [f"cuda:{i}" for i in range(len(self._devices))]
        else:
            self._devices = list(self._devices)

    def _tie_width(self, N: int, dtype, kwargs) -> int:
        """The base solver's suffix width B on the whole problem (its tie rule)."""
        if self._base_solver is QB_F1:
            return min(QB_F1.SUFFIX_BITS, N)
        kw = {k: v for k, v in kwargs.items() if k in ("A", "B", "use_mem_GB")}
        kw.setdefault("size_of", np.dtype(dtype).itemsize)
        A, B = define_chunk_size(N, **kw)
        if not (0 <= B <= N):
            raise ValueError(f"B={B} outside [0, N={N}]")
        return B

    def _metrics(self, tasks, n_sub: int, kwargs) -> dict:
        """The reference's merged metrics (parallel.py:245-267): every task's energy_arr
        (plus its prefix_E) and time_arr, concatenated in task order.  A task's metrics are
        those its own base solve would record: per prefix of the reduced problem under the
        sub-solve's own A/B split (the kwargs applied to N - k variables), from one walk of
        the reduced matrix (on the first device)."""
        use = QBF.metrics_wanted(kwargs)
        if not use or n_sub == 0:
            return {}
        kw = {k: v for k, v in kwargs.items() if k in ("A", "B", "use_mem_GB")}
        kw.setdefault("size_of", 8)
        _, B_sub = define_chunk_size(n_sub, **kw)
        if self._base_solver is QB_F1:
            B_sub = min(QB_F1.SUFFIX_BITS, n_sub)
        if not (1 <= B_sub <= n_sub):
            raise ValueError(f"use_metrics needs 1 <= B <= {n_sub} for the reduced problems (B={B_sub})")
        dev = _native.parse_device(self._devices[0])
        energy, times = [], []
        for R, pe, _, _, R0 in tasks:
            _, minima, wave_ms = engine.solve_matrix_minima(R0, device=dev, tie_B=B_sub)
            energy.append(np.asarray(minima, dtype=np.float64) + float(pe))
            times.append(wave_ms)
        out = {}
        if "time_arr" in use:
            out["time_arr"] = np.concatenate(times)
        if "energy_arr" in use:
            out["energy_arr"] = np.concatenate(energy)
        return out

    def solve(self, problem, timeout=None, **kwargs) -> Solution:
        SolverABC.check_obsolete_args(**kwargs)
        kw = dict(solver_type=self._solver_class_name, splits_cnt=self._splits_cnt,
                  devices=self._devices, timeout=timeout, **kwargs)
        json_str = self.encode_kwargs(kw)
        N = problem.N
        k = self._splits_cnt.bit_length() - 1
        if k > N:
            raise ValueError(f"splits_cnt=2^{k} exceeds 2^N for N={N}")
        if self._solver_class_name == "QB_F1" and np.dtype(problem.dtype) != np.int16:
            raise TypeError("Q dtype only np.int16 supported yet")
        Q = np.asarray(problem.build_Q_matrix(use_gpu=False))
        B = self._tie_width(N, Q.dtype, kwargs)
        A = N - B
        n_sub = N - k
        # the tie width of every sub-solve, and whether odd-parity tasks flip their first bit
        sub_B = n_sub if k >= A else B
        flip_odd = k < A and n_sub > 0
        tasks = []
        for task_id in range(self._splits_cnt):
            prefix = np.array([(task_id >> (k - 1 - j)) & 1 for j in range(k)], dtype=np.int64)
            R0, pe = reduced_matrix(Q, prefix)
            flip = flip_odd and bin(task_id).count("1") % 2 == 1
            R, offset = complement_first(R0) if flip else (R0, 0)
            tasks.append((R, pe, flip, offset, R0))
        ndev = len(self._devices)
        shares = [list(range(d, self._splits_cnt, ndev)) for d in range(ndev)]
        results: dict = {}
        errors: list = []

        def work(worker_id):
            dev = _native.parse_device(self._devices[worker_id])
            ids = shares[worker_id]
            if not ids:
                return
            t0 = time.perf_counter()
            try:
                if n_sub == 0:
                    raws = [None] * len(ids)
                else:
                    raws = engine.solve_matrices([tasks[t][0] for t in ids], device=dev, tie_B=sub_B,
                                                 walk_bits=kwargs.get("walk_bits"),
                                                 inner_bits=kwargs.get("inner_bits"))
                ms = (time.perf_counter() - t0) * 1000.0
                for t, r in zip(ids, raws):
                    results[t] = (worker_id, r, ms / len(ids))
            except Exception as e:  # reported on the merged Solution, never silently dropped
                errors.append((worker_id, e))

        t_start = time.perf_counter()
        if self._devices[0] == "cpu":
            for w in range(ndev):          # one host backend: shares one after another
                work(w)
        else:
            threads = [threading.Thread(target=work, args=(w,)) for w in range(ndev)]
            for t in threads:
                t.start()
            for t in threads:
                t.join()
        wall_ms = (time.perf_counter() - t_start) * 1000.0
        if errors:
            raise errors[0][1]

        merged = Solution(N=N, epochs_cnt=self._splits_cnt, samples_cnt=1,
                          problem_instance_json=problem.to_json(), solver_params_json=json_str)
        merged.ising_offset = None
        merged.solver_class = self._solver_class_name
        best = None
        integer = np.issubdtype(Q.dtype, np.integer)
        for task_id in range(self._splits_cnt):
            _, r, _ = results[task_id]
            R, pe, flip, offset, _ = tasks[task_id]
            if r is None:
                e_sub, s_sub = 0, 0
            else:
                e_sub = r.E + offset
                s_sub = r.state ^ ((1 << (n_sub - 1)) if flip else 0)
            total = e_sub + pe
            if not integer:
                total = float(total)
            state = (task_id << n_sub) | s_sub
            key = (total, full_rank(state, B))
            if best is None or key < best[0]:
                best = (key, state)
        (total, _), state = best
        full = engine.state_bits(state, N)
        merged.best_E = np.float64(total)      # a float64, as QBF's (qbf.py:356-360)
        merged.best_E_exact = total
        merged.energies = np.array([np.float64(total)])
        merged.best_X = full
        merged.state_vectors = StateVectors(wide_vector_matrix=full.reshape(1, -1))
        merged.metrics = [self._metrics(tasks, n_sub, kwargs)]
        worker_times: dict = {}
        for t, (w, _, ms) in results.items():
            worker_times[w] = worker_times.get(w, 0.0) + ms
        merged.time_to_solve_ms = max(worker_times.values()) if worker_times else 0.0
        merged.worker_logs = [dict(task_id=t, worker_id=w, device=self._devices[w], time_to_solve_ms=ms)
                              for t, (w, _, ms) in sorted(results.items())]
        merged.tie_B = B
        if timeout is not None and wall_ms > timeout * 1000.0:
            merged.error = f"timeout: solve took {wall_ms:.1f} ms > {timeout} s"
        return merged
