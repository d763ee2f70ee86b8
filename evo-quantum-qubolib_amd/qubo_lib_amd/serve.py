"""Continuous-batching server: solve requests as they arrive, sharing the GPU.

    with Server() as srv:
        rid = srv.submit(problem_or_Q)
        ...
        for res in srv.poll(wait_ms=10): ...

The native worker (libqbx.so, qbx_server_*) admits requests into device slots while the
in-flight work (sum of 2^n) fits a budget and runs scheduling steps -- finite persistent
launches over chunks of all admitted requests -- so requests arriving while others run
are solved alongside them.  Every answer equals the one `QBF().solve` / `engine.solve_matrix`
gives for the same matrix (same kernels, same tie rule).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native, engine

_P = ctypes.POINTER
class _Packed:
    """What poll needs of a request submitted packed: its path and value type."""
    __slots__ = ("kind", "integer", "unit")

    def __init__(self, kind, integer):
        self.kind, self.integer, self.unit = kind, integer, 1.0


_RESULT_FIELDS = tuple(name for name, _ in _native.RequestResult._fields_)
_RESULT_DTYPE = np.dtype({"names": list(_RESULT_FIELDS),
                          "formats": [np.dtype(t) for _, t in _native.RequestResult._fields_],
                          "offsets": [getattr(_native.RequestResult, f).offset for f in _RESULT_FIELDS],
                          "itemsize": ctypes.sizeof(_native.RequestResult)})


class Server:
    def __init__(self, device: int = 0, max_n: int = 40, max_inflight: int = 512,
                 work_budget: float = 2.0 ** 40, step_states: float = 2.0 ** 34, walk_bits=None):
        self.ctx = _native.get_context(device)
        self.lib = self.ctx.lib
        o = _native.ServerOpts(int(max_n), int(max_inflight), float(work_budget), float(step_states),
                               -1 if walk_bits is None else int(walk_bits), 0)
        self._h = ctypes.c_void_p()
        _native._check(self.lib, self.lib.qbx_server_start(self.ctx._h, ctypes.byref(o), ctypes.byref(self._h)))
        self.max_n = int(max_n)
        self._kinds = {}
        self._buf = (_native.RequestResult * 1024)()
        # the same buffer as records, to read a whole poll at once
        self._view = np.frombuffer(self._buf, dtype=_RESULT_DTYPE)

    def submit(self, problem, tie_B=None) -> int:
        """Queue one problem (a ProblemABC or a square matrix); returns its request id."""
        Q = problem if isinstance(problem, (np.ndarray, engine.Prepared)) else problem.build_Q_matrix(use_gpu=False)
        p = Q if isinstance(Q, engine.Prepared) else engine.prepare(Q)
        if p.n > self.max_n:
            raise ValueError(f"N={p.n} exceeds the server's max_n={self.max_n}")
        rid = ctypes.c_uint64()
        tb = -1 if tie_B is None else int(tie_B)
        if p.kind == "i32":
            m = np.ascontiguousarray(p.matrix, dtype=np.int32)
            rc = self.lib.qbx_server_submit_i32(self._h, m.ctypes.data_as(_P(ctypes.c_int32)), p.n, tb, ctypes.byref(rid))
        else:
            m = np.ascontiguousarray(p.matrix, dtype=np.float64)
            rc = self.lib.qbx_server_submit_f64(self._h, m.ctypes.data_as(_P(ctypes.c_double)), p.n, tb, ctypes.byref(rid))
        _native._check(self.lib, rc)
        self._kinds[rid.value] = p
        return rid.value

    SUBMIT_CHUNK = 2048

    def submit_many(self, problems, tie_B=None) -> list:
        """Queue a burst of problems; returns their request ids in input order.  Every
        problem is checked first (a bad one queues none); they are then packed and handed to
        the library SUBMIT_CHUNK at a time, so the GPU starts on the first ones while the
        rest are packed."""
        preps = []
        for problem in problems:
            Q = problem if isinstance(problem, (np.ndarray, engine.Prepared)) else problem.build_Q_matrix(use_gpu=False)
            p = Q if isinstance(Q, engine.Prepared) else engine.prepare(Q)
            if p.n > self.max_n:
                raise ValueError(f"N={p.n} exceeds the server's max_n={self.max_n}")
            preps.append(p)
        tb = -1 if tie_B is None else int(tie_B)
        ids = [0] * len(preps)
        for kind, dtype, fn, ptr in (("i32", np.int32, self.lib.qbx_server_submit_many_i32, ctypes.c_int32),
                                     ("f64", np.float64, self.lib.qbx_server_submit_many_f64, ctypes.c_double)):
            idx_all = [i for i, p in enumerate(preps) if p.kind == kind]
            for c0 in range(0, len(idx_all), self.SUBMIT_CHUNK):
                idx = idx_all[c0:c0 + self.SUBMIT_CHUNK]
                data, offsets, ns = _native.Context.pack([preps[i].matrix for i in idx], dtype)
                out = np.zeros(len(idx), dtype=np.uint64)
                _native._check(self.lib, fn(self._h, len(idx), data.ctypes.data_as(_P(ptr)),
                                            offsets.ctypes.data_as(_P(ctypes.c_int64)),
                                            ns.ctypes.data_as(_P(ctypes.c_int32)), tb,
                                            out.ctypes.data_as(_P(ctypes.c_uint64))))
                for i, rid in zip(idx, out.tolist()):
                    ids[i] = rid
                    self._kinds[rid] = preps[i]
        return ids

    def submit_packed(self, data, offsets, ns, kind: str = "i32", tie_B=None, integer: bool = True) -> np.ndarray:
        """Queue a burst already packed by the caller (Context.pack layout: matrices of one
        arithmetic path back to back; folded int32 for "i32", float64 for "f64"); returns
        the request ids in order.  No per-request Python work: for clients that keep their
        requests packed.  `integer`: the matrices' values are integers (energies come back
        as ints)."""
        dtype, fn, ptr = ((np.int32, self.lib.qbx_server_submit_many_i32, ctypes.c_int32) if kind == "i32"
                          else (np.float64, self.lib.qbx_server_submit_many_f64, ctypes.c_double))
        data = np.ascontiguousarray(data, dtype=dtype)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        ns = np.ascontiguousarray(ns, dtype=np.int32)
        if ns.size and int(ns.max()) > self.max_n:
            raise ValueError(f"N={int(ns.max())} exceeds the server's max_n={self.max_n}")
        out = np.zeros(ns.size, dtype=np.uint64)
        tb = -1 if tie_B is None else int(tie_B)
        _native._check(self.lib, fn(self._h, int(ns.size), data.ctypes.data_as(_P(ptr)),
                                    offsets.ctypes.data_as(_P(ctypes.c_int64)), ns.ctypes.data_as(_P(ctypes.c_int32)),
                                    tb, out.ctypes.data_as(_P(ctypes.c_uint64))))
        tag = _Packed(kind, integer)
        self._kinds.update(dict.fromkeys(out.tolist(), tag))
        return out

    def poll(self, wait_ms: int = 0, max_results: int = 1024) -> list:
        """Completed requests as dicts: id, E (int / float as the problem), state, rank,
        status, n and the server-clock times t_submit_ms / t_admit_ms / t_done_ms."""
        cnt = ctypes.c_int()
        _native._check(self.lib, self.lib.qbx_server_poll(self._h, self._buf, min(max_results, 1024),
                                                          int(wait_ms), ctypes.byref(cnt)))
        out = []
        for rec in self._view[:cnt.value].tolist():
            r = dict(zip(_RESULT_FIELDS, rec))
            p = self._kinds.pop(r["id"], None)
            if p is not None:
                if p.integer:
                    r["E"] = int(r["E_int"]) if p.kind == "i32" else int(round(r["E"]))
                elif p.kind == "i32":
                    r["E"] = engine.energy_of(p, r["E_int"])
            out.append(r)
        return out

    def stats(self) -> dict:
        s = _native.ServerStats()
        _native._check(self.lib, self.lib.qbx_server_stats(self._h, ctypes.byref(s)))
        return {name: getattr(s, name) for name, _ in _native.ServerStats._fields_}

    def close(self):
        if self._h:
            h, self._h = self._h, ctypes.c_void_p()
            _native._check(self.lib, self.lib.qbx_server_stop(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
