"""CPU oracle for the exhaustive QUBO path -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
always as the checker.  The product package (qubo_lib_amd) never imports it.

  qbf_i64 / qbf_f64     C restatement of the reference QBF algorithm (qbf_oracle.c)
  brute_i64 / brute_f64 independent C brute force, exact (E, rank) tie rule
  qbf_numpy             numpy restatement of QBF (qbf_numpy.py), the timed CPU baseline
  energy                direct x^T Q x
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from .build import OUT as LIB_PATH, build

_P = ctypes.POINTER
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        for name in ("oracle_qbf_i64", "oracle_brute_i64"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [_P(ctypes.c_int64), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          _P(ctypes.c_int64), _P(ctypes.c_uint64)]
        for name in ("oracle_qbf_f64", "oracle_brute_f64"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [_P(ctypes.c_double), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          _P(ctypes.c_double), _P(ctypes.c_uint64)]
        L.oracle_rank_of_state.restype = ctypes.c_uint64
        L.oracle_rank_of_state.argtypes = [ctypes.c_uint64, ctypes.c_int]
        L.oracle_energy_i64.restype = ctypes.c_int64
        L.oracle_energy_i64.argtypes = [_P(ctypes.c_int64), ctypes.c_int, ctypes.c_uint64]
        L.oracle_energy_f64.restype = ctypes.c_double
        L.oracle_energy_f64.argtypes = [_P(ctypes.c_double), ctypes.c_int, ctypes.c_uint64]
        _lib = L
    return _lib


def _default_B(n, B):
    return min(n, 20) if B is None else int(B)


def _call_i(fn, Q, B, extra):
    Q = np.ascontiguousarray(Q, dtype=np.int64)
    n = Q.shape[0]
    E = ctypes.c_int64()
    x = ctypes.c_uint64()
    rc = fn(Q.ctypes.data_as(_P(ctypes.c_int64)), n, _default_B(n, B), extra, ctypes.byref(E), ctypes.byref(x))
    if rc != 0:
        raise ValueError(f"oracle rejected n={n}, B={B}")
    return int(E.value), int(x.value)


def _call_f(fn, Q, B, extra):
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    n = Q.shape[0]
    E = ctypes.c_double()
    x = ctypes.c_uint64()
    rc = fn(Q.ctypes.data_as(_P(ctypes.c_double)), n, _default_B(n, B), extra, ctypes.byref(E), ctypes.byref(x))
    if rc != 0:
        raise ValueError(f"oracle rejected n={n}, B={B}")
    return float(E.value), int(x.value)


def qbf_i64(Q, B=None, correction_period=1024):
    """Restatement of QBF (int64 accumulation): returns (E, state)."""
    return _call_i(lib().oracle_qbf_i64, Q, B, correction_period)


def qbf_f64(Q, B=None, correction_period=1024):
    return _call_f(lib().oracle_qbf_f64, Q, B, correction_period)


def brute_i64(Q, B=None, threads=0):
    """Independent brute force, exact (E, rank) minimum: returns (E, state)."""
    return _call_i(lib().oracle_brute_i64, Q, B, threads)


def brute_f64(Q, B=None, threads=0):
    return _call_f(lib().oracle_brute_f64, Q, B, threads)


def rank_of_state(s, B):
    return int(lib().oracle_rank_of_state(int(s), int(B)))


def energy(Q, s):
    Q = np.asarray(Q)
    n = Q.shape[0]
    if np.issubdtype(Q.dtype, np.integer):
        Qc = np.ascontiguousarray(Q, dtype=np.int64)
        return int(lib().oracle_energy_i64(Qc.ctypes.data_as(_P(ctypes.c_int64)), n, int(s)))
    Qc = np.ascontiguousarray(Q, dtype=np.float64)
    return float(lib().oracle_energy_f64(Qc.ctypes.data_as(_P(ctypes.c_double)), n, int(s)))
