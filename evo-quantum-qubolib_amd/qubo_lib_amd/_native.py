"""ctypes bindings of libqbx.so, the gfx950 exhaustive-QUBO library (C ABI in include/qbx.h),
and of libqbx_cpu.so, the same lane walk built for the host (include/qbx_cpu.h), which
serves device='cpu'.

This module is the only place the product touches native code.  A 'cuda' device never
falls back to the CPU: if libqbx.so is missing or no GPU is visible, its entry points raise.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

import numpy as np

__all__ = [
    "LIB_PATH", "NativeError", "Context", "Opts", "Result", "DevInfo", "ServerOpts", "RequestResult",
    "ServerStats", "Pcg64State", "pcg64_states", "GEN_DTYPES", "BatchPlan", "KINDS", "MODES", "load_library", "get_context", "read_tuning",
    "tuning_path", "PATHS", "PAIR_BOUND", "CpuBackend", "load_cpu_library", "get_backend", "parse_device",
]

LIB_PATH = Path(__file__).resolve().parent / "libqbx.so"
CPU_LIB_PATH = Path(__file__).resolve().parent / "libqbx_cpu.so"
TUNING_DIR = Path(__file__).resolve().parent / "tuning"

QBX_OK = 0
QBX_EINVAL = -1
QBX_ERANGE = -2
QBX_EDEVICE = -3
QBX_ENOMEM = -4
QBX_EBUSY = -5

KINDS = {"i32": 0, "f64": 1}
MODES = {"single": 0, "batch": 1}


class NativeError(RuntimeError):
    """A libqbx.so call failed (device, memory or range error)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"libqbx error {code}: {message}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [
        ("tie_B", ctypes.c_int32),
        ("walk_bits", ctypes.c_int32),
        ("inner_bits", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


# qbx_result_t.path: the walk that ran (include/qbx.h QBX_PATH_*)
PATHS = {0: "int32", 1: "pair16", 2: "float", 3: "pair16_rebased", 4: "float_pair16"}
PAIR_BOUND = 15359       # QBX_PAIR_BOUND: integer bound up to which the packed pair walk runs


class Result(ctypes.Structure):
    _fields_ = [
        ("E_int", ctypes.c_int64),
        ("E", ctypes.c_double),
        ("state", ctypes.c_uint64),
        ("rank", ctypes.c_uint64),
        ("states", ctypes.c_uint64),
        ("tiles", ctypes.c_uint64),
        ("kernel_ms", ctypes.c_double),
        ("n", ctypes.c_int32),
        ("tie_B", ctypes.c_int32),
        ("walk_bits", ctypes.c_int32),
        ("inner_bits", ctypes.c_int32),
        ("lane_bits", ctypes.c_int32),
        ("status", ctypes.c_int32),
        ("path", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class DevInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int32),
        ("clock_khz", ctypes.c_int32),
        ("simd_per_cu", ctypes.c_int32),
        ("wave_size", ctypes.c_int32),
        ("total_mem", ctypes.c_int64),
        ("valu_peak_ops", ctypes.c_double),
    ]


class Pcg64State(ctypes.Structure):
    _fields_ = [
        ("state_hi", ctypes.c_uint64),
        ("state_lo", ctypes.c_uint64),
        ("inc_hi", ctypes.c_uint64),
        ("inc_lo", ctypes.c_uint64),
    ]


def pcg64_states(seeds, lib=None):
    """The states numpy's Generator(PCG64(seed)) starts from: computed by the library
    (qbx_pcg64_from_seeds) for integer seeds in [0, 2^64), else by numpy's SeedSequence."""
    states = (Pcg64State * len(seeds))()
    if all(isinstance(s, (int, np.integer)) and 0 <= int(s) < 2 ** 64 for s in seeds):
        lib = lib or load_library()
        arr = np.ascontiguousarray(np.asarray([int(s) for s in seeds], dtype=np.uint64))
        _check(lib, lib.qbx_pcg64_from_seeds(len(seeds), arr.ctypes.data_as(_P(ctypes.c_uint64)), states))
        return states
    from numpy.random import PCG64
    for st, seed in zip(states, seeds):
        s = PCG64(seed).state["state"]
        st.state_hi, st.state_lo = s["state"] >> 64, s["state"] & ((1 << 64) - 1)
        st.inc_hi, st.inc_lo = s["inc"] >> 64, s["inc"] & ((1 << 64) - 1)
    return states


GEN_DTYPES = {"int16": 16, "int32": 32, "float32": -32, "float64": -64}


class ServerOpts(ctypes.Structure):
    _fields_ = [
        ("max_n", ctypes.c_int32),
        ("max_inflight", ctypes.c_int32),
        ("work_budget", ctypes.c_double),
        ("step_states", ctypes.c_double),
        ("walk_bits", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


class RequestResult(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("E_int", ctypes.c_int64),
        ("E", ctypes.c_double),
        ("state", ctypes.c_uint64),
        ("rank", ctypes.c_uint64),
        ("t_submit_ms", ctypes.c_double),
        ("t_admit_ms", ctypes.c_double),
        ("t_done_ms", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("steps", ctypes.c_uint64),
        ("chunks", ctypes.c_uint64),
        ("admissions", ctypes.c_uint64),
        ("busy_ms", ctypes.c_double),
    ]


# every symbol declared in include/qbx.h, with its ctypes signature
_P = ctypes.POINTER
_SIGNATURES = {
    "qbx_open": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_void_p)]),
    "qbx_close": (ctypes.c_int, [ctypes.c_void_p]),
    "qbx_last_error": (ctypes.c_char_p, []),
    "qbx_device_info": (ctypes.c_int, [ctypes.c_void_p, _P(DevInfo)]),
    "qbx_solve_i32": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), ctypes.c_int, _P(Opts), _P(Result)]),
    "qbx_solve_f64": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double), ctypes.c_int, _P(Opts), _P(Result)]),
    "qbx_solve_i32_noalloc": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), ctypes.c_int, _P(Opts),
                                             _P(Result)]),
    "qbx_solve_minima_i32": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), ctypes.c_int, _P(Opts),
                                            _P(ctypes.c_int64), _P(ctypes.c_double), _P(Result)]),
    "qbx_solve_minima_f64": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double), ctypes.c_int, _P(Opts),
                                            _P(ctypes.c_double), _P(ctypes.c_double), _P(Result)]),
    "qbx_solve_part_i32": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), ctypes.c_int, _P(Opts), ctypes.c_int,
                                          ctypes.c_int, _P(Result)]),
    "qbx_solve_part_f64": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double), ctypes.c_int, _P(Opts), ctypes.c_int,
                                          ctypes.c_int, _P(Result)]),
    "qbx_batch_i32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_int32), _P(ctypes.c_int64),
                                     _P(ctypes.c_int32), _P(Opts), _P(Result)]),
    "qbx_batch_f64": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_int64),
                                     _P(ctypes.c_int32), _P(Opts), _P(Result)]),
    "qbx_pcg64_from_seeds": (ctypes.c_int, [ctypes.c_int, _P(ctypes.c_uint64), _P(Pcg64State)]),
    "qbx_batch_random_qubo": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(Pcg64State), _P(ctypes.c_int32),
                                             ctypes.c_int64, ctypes.c_int64, ctypes.c_int, _P(Opts), _P(Result)]),
    "qbx_random_qubo_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(Pcg64State), _P(ctypes.c_int32),
                                            ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p]),
    "qbx_batch_random_tsp": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(Pcg64State), _P(ctypes.c_int32),
                                            ctypes.c_double, ctypes.c_double, ctypes.c_int, _P(Opts), _P(Result)]),
    "qbx_random_tsp_read": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(Pcg64State), _P(ctypes.c_int32),
                                           ctypes.c_double, ctypes.c_double, ctypes.c_int, _P(ctypes.c_double)]),
    "qbx_set_walk_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int32),
                                          ctypes.c_int]),
    "qbx_plan_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P(Opts),
                                     _P(Result)]),
    "qbx_batch_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _P(ctypes.c_int32),
                                             _P(Opts), _P(ctypes.c_void_p)]),
    "qbx_batch_plan_run_i32": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), _P(ctypes.c_int64),
                                              _P(Result)]),
    "qbx_batch_plan_run_f64": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double), _P(ctypes.c_int64),
                                              _P(Result)]),
    "qbx_batch_plan_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "qbx_server_start": (ctypes.c_int, [ctypes.c_void_p, _P(ServerOpts), _P(ctypes.c_void_p)]),
    "qbx_server_submit_i32": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_int32), ctypes.c_int, ctypes.c_int,
                                             _P(ctypes.c_uint64)]),
    "qbx_server_submit_f64": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_double), ctypes.c_int, ctypes.c_int,
                                             _P(ctypes.c_uint64)]),
    "qbx_server_submit_many_i32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_int32), _P(ctypes.c_int64),
                                                  _P(ctypes.c_int32), ctypes.c_int, _P(ctypes.c_uint64)]),
    "qbx_server_submit_many_f64": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_int64),
                                                  _P(ctypes.c_int32), ctypes.c_int, _P(ctypes.c_uint64)]),
    "qbx_server_poll": (ctypes.c_int, [ctypes.c_void_p, _P(RequestResult), ctypes.c_int, ctypes.c_int,
                                       _P(ctypes.c_int)]),
    "qbx_server_stats": (ctypes.c_int, [ctypes.c_void_p, _P(ServerStats)]),
    "qbx_server_stop": (ctypes.c_int, [ctypes.c_void_p]),
}
EXPORTED_SYMBOLS = tuple(_SIGNATURES)

_lib = None
_lib_lock = threading.Lock()


def load_library(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libqbx.so (built in-tree by __graft_entry__.build()).  Raises if it is missing."""
    global _lib
    with _lib_lock:
        if _lib is not None and path is None:
            return _lib
        p = Path(path) if path is not None else LIB_PATH
        if not p.exists():
            raise NativeError(QBX_EDEVICE, f"{p} not found: build it with __graft_entry__.build()")
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def tuning_path(arch: str, path=None) -> Path | None:
    """The tiling table to load: `path`, else $QUBO_LIB_AMD_TUNING ("off": none), else
    tuning/<arch>.json next to this module; None when the file does not exist."""
    env = os.environ.get("QUBO_LIB_AMD_TUNING")
    if path is None and env:
        if env.lower() == "off":
            return None
        path = env
    if path is None:
        path = TUNING_DIR / f"{arch}.json"
    path = Path(path)
    return path if path.exists() else None


def read_tuning(path) -> dict:
    """{(kind, mode): walk bits per n} from a table written by qubo_lib_amd.tune.  Raises
    ValueError on an unknown kind/mode or an entry outside -1 .. n."""
    import json
    data = json.loads(Path(path).read_text())
    out = {}
    for key, table in data.get("walk_bits", {}).items():
        kind, _, mode = key.partition("/")
        if kind not in KINDS or mode not in MODES:
            raise ValueError(f"{path}: unknown table {key!r}")
        table = [int(v) for v in table]
        if len(table) > 57 or any(v < -1 or v > n for n, v in enumerate(table)):
            raise ValueError(f"{path}: table {key!r} has an entry outside -1..n")
        out[(kind, mode)] = table
    return out


def _check(lib, rc: int):
    if rc == QBX_OK:
        return
    msg = lib.qbx_last_error().decode(errors="replace")
    if rc == QBX_EINVAL:
        raise ValueError(msg)
    if rc == QBX_ERANGE:
        raise OverflowError(msg)
    raise NativeError(rc, msg)


# float problems: None = by size (packed from n = 22 in batches, n = 28 alone), or forced
FLOAT_FILTERS = {None: 0, "int32": 1, "packed": 2}   # QBX_FLAG_FLOAT_INT32_FILTER / _PACKED_FILTER


MAX_PREFIX_BITS = 28     # QBX_MAX_PREFIX_BITS: per-prefix records of at most 2^28 prefixes


def _minima_buffers(n: int, kind: str, kw):
    """Output arrays of a per-prefix solve (2^(n - B) entries) and its options."""
    B = min(n, 20) if kw.get("tie_B") is None else int(kw["tie_B"])
    if not (1 <= B <= n) or n - B > MAX_PREFIX_BITS:
        raise ValueError(f"per-prefix records need 1 <= B <= N and N - B <= {MAX_PREFIX_BITS} (N={n}, B={B})")
    count = 1 << (n - B)
    minima = np.empty(count, dtype=np.int64 if kind == "i32" else np.float64)
    return minima, np.empty(count, dtype=np.float64), _opts(**kw)


def _opts(tie_B=None, walk_bits=None, inner_bits=None, float_filter=None) -> Opts:
    """float_filter: None (by problem size), "packed" (wherever it can run) or "int32"."""
    if float_filter not in FLOAT_FILTERS:
        raise ValueError(f"float_filter must be one of {sorted(k for k in FLOAT_FILTERS if k)}: {float_filter!r}")
    return Opts(-1 if tie_B is None else int(tie_B),
                -1 if walk_bits is None else int(walk_bits),
                -1 if inner_bits is None else int(inner_bits), FLOAT_FILTERS[float_filter])


class Context:
    """One HIP device and stream.  Not thread-safe: one host thread at a time."""

    def __init__(self, device: int = 0, tuning: str | os.PathLike | None = None):
        self.lib = load_library()
        self._h = ctypes.c_void_p()
        _check(self.lib, self.lib.qbx_open(int(device), ctypes.byref(self._h)))
        self.device = int(device)
        self.lock = threading.Lock()
        self.tuning_file = None
        self.load_tuning(tuning)

    # -- tiling -----------------------------------------------------------------------
    def load_tuning(self, path=None) -> Path | None:
        """Install the per-n walk-bit tables written by qubo_lib_amd.tune for this device's
        architecture.  Looked up in order: `path`, $QUBO_LIB_AMD_TUNING ("off" disables),
        qubo_lib_amd/tuning/<arch>.json.  A missing file keeps the built-in rule."""
        path = tuning_path(self.device_info()["name"].split(":")[0], path)
        if path is None:
            return None
        for (kind, mode), table in read_tuning(path).items():
            self.set_walk_table(kind, mode, table)
        self.tuning_file = path
        return path

    def set_walk_table(self, kind: str, mode: str, table) -> None:
        """walk_bits per n (index = n, -1 = built-in rule) for one kind ("i32"/"f64") and
        mode ("single"/"batch")."""
        arr = np.ascontiguousarray(np.asarray(table, dtype=np.int32))
        with self.lock:
            rc = self.lib.qbx_set_walk_table(self._h, KINDS[kind], MODES[mode],
                                             arr.ctypes.data_as(_P(ctypes.c_int32)), int(arr.size))
        _check(self.lib, rc)

    def plan_info(self, kind: str, mode: str, n: int, **kw) -> Result:
        """Tiling a solve of size n would get (walk_bits, inner_bits, lane_bits, tiles)."""
        out = Result()
        o = _opts(**kw)
        with self.lock:
            rc = self.lib.qbx_plan_info(self._h, KINDS[kind], MODES[mode], int(n), ctypes.byref(o),
                                        ctypes.byref(out))
        _check(self.lib, rc)
        return out

    def close(self):
        if self._h:
            self.lib.qbx_close(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self) -> dict:
        info = DevInfo()
        _check(self.lib, self.lib.qbx_device_info(self._h, ctypes.byref(info)))
        d = {name: getattr(info, name) for name, _ in DevInfo._fields_}
        d["name"] = info.name.decode()
        return d

    # -- single problems ------------------------------------------------------------
    def solve_i32(self, Q: np.ndarray, **kw) -> Result:
        Q = np.ascontiguousarray(Q, dtype=np.int32)
        n = Q.shape[0]
        out = Result()
        o = _opts(**kw)
        with self.lock:
            rc = self.lib.qbx_solve_i32(self._h, Q.ctypes.data_as(_P(ctypes.c_int32)), n,
                                        ctypes.byref(o), ctypes.byref(out))
        _check(self.lib, rc)
        return out

    def solve_f64(self, Q: np.ndarray, **kw) -> Result:
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        n = Q.shape[0]
        out = Result()
        o = _opts(**kw)
        with self.lock:
            rc = self.lib.qbx_solve_f64(self._h, Q.ctypes.data_as(_P(ctypes.c_double)), n,
                                        ctypes.byref(o), ctypes.byref(out))
        _check(self.lib, rc)
        return out

    def solve_i32_noalloc(self, Q: np.ndarray, **kw) -> Result:
        """qbx_solve_i32 in the library's fixed device scratch (no allocation; calls that
        overlap on one GPU race)."""
        Q = np.ascontiguousarray(Q, dtype=np.int32)
        out = Result()
        o = _opts(**kw)
        with self.lock:
            rc = self.lib.qbx_solve_i32_noalloc(self._h, Q.ctypes.data_as(_P(ctypes.c_int32)), Q.shape[0],
                                                ctypes.byref(o), ctypes.byref(out))
        _check(self.lib, rc)
        return out

    def solve_minima(self, Q: np.ndarray, kind: str = "i32", **kw):
        """(Result, minima, wave_ms): the solve plus, per reference prefix in Gray visiting
        order, its least energy (int64 / float64) and the device time spent on it (ms)."""
        dtype, fn, ptr = ((np.int32, self.lib.qbx_solve_minima_i32, ctypes.c_int32) if kind == "i32"
                          else (np.float64, self.lib.qbx_solve_minima_f64, ctypes.c_double))
        Q = np.ascontiguousarray(Q, dtype=dtype)
        minima, wave_ms, o = _minima_buffers(Q.shape[0], kind, kw)
        out = Result()
        with self.lock:
            rc = fn(self._h, Q.ctypes.data_as(_P(ptr)), Q.shape[0], ctypes.byref(o),
                    minima.ctypes.data_as(_P(ctypes.c_int64 if kind == "i32" else ctypes.c_double)),
                    wave_ms.ctypes.data_as(_P(ctypes.c_double)), ctypes.byref(out))
        _check(self.lib, rc)
        return out, minima, wave_ms

    def solve_part(self, Q: np.ndarray, part: int, parts: int, kind: str = "i32", **kw) -> Result:
        """(E, rank) minimum over part `part` of `parts` equal ranges of the solve's chunks
        (one GPU's share of a solve spread over `parts` GPUs)."""
        dtype, fn, ptr = ((np.int32, self.lib.qbx_solve_part_i32, ctypes.c_int32) if kind == "i32"
                          else (np.float64, self.lib.qbx_solve_part_f64, ctypes.c_double))
        Q = np.ascontiguousarray(Q, dtype=dtype)
        out = Result()
        o = _opts(**kw)
        with self.lock:
            rc = fn(self._h, Q.ctypes.data_as(_P(ptr)), Q.shape[0], ctypes.byref(o), int(part), int(parts),
                    ctypes.byref(out))
        _check(self.lib, rc)
        return out

    # -- batches --------------------------------------------------------------------
    @staticmethod
    def pack(Qs, dtype):
        """(packed data, offsets, ns) of square matrices, row-major, back to back."""
        ns = np.array([q.shape[0] for q in Qs], dtype=np.int32)
        sizes = ns.astype(np.int64) ** 2
        offsets = np.zeros(len(Qs), dtype=np.int64)
        if len(Qs) > 1:
            offsets[1:] = np.cumsum(sizes)[:-1]
        if not len(Qs):
            return np.empty(0, dtype=dtype), offsets, ns
        packed = np.concatenate([np.asarray(q).reshape(-1) for q in Qs]).astype(dtype, copy=False)
        return np.ascontiguousarray(packed), offsets, ns

    def batch(self, Qs, kind: str = "i32", packed=None, **kw):
        """Solve a list of problems in one persistent launch; returns a sequence of Result
        (the ctypes array the library filled: indexing and iteration give views of it, so
        no per-result Python objects are made until they are used)."""
        dtype, fn, ptr = ((np.int32, self.lib.qbx_batch_i32, ctypes.c_int32) if kind == "i32"
                          else (np.float64, self.lib.qbx_batch_f64, ctypes.c_double))
        if packed is None:
            packed = self.pack(Qs, dtype)
        data, offsets, ns = packed
        count = len(ns)
        out = (Result * count)()
        o = _opts(**kw)
        with self.lock:
            rc = fn(self._h, count, data.ctypes.data_as(_P(ptr)), offsets.ctypes.data_as(_P(ctypes.c_int64)),
                    ns.ctypes.data_as(_P(ctypes.c_int32)), ctypes.byref(o), out)
        if rc not in (QBX_OK, QBX_ERANGE):
            _check(self.lib, rc)
        return out

    # -- RandomQUBO batches drawn on the device ----------------------------------------
    def _gen_args(self, seeds, ns, dtype):
        states = pcg64_states(seeds, self.lib)
        nsa = np.ascontiguousarray(np.asarray(ns, dtype=np.int32))
        if nsa.size != len(seeds):
            raise ValueError("one N per seed")
        key = np.dtype(dtype).name
        if key not in GEN_DTYPES:
            raise TypeError(f"RandomQUBO dtype {key} cannot be drawn on the device")
        return states, nsa, GEN_DTYPES[key]

    def random_qubo_batch(self, seeds, ns, low=-100, high=100, dtype=np.float32, **kw):
        """Solve RandomQUBO(seed, N, low, high, dtype) instances whose matrices are drawn on
        the device (bit for bit numpy's); returns a sequence of Result (as batch())."""
        states, nsa, code = self._gen_args(seeds, ns, dtype)
        out = (Result * len(seeds))()
        o = _opts(**kw)
        with self.lock:
            rc = self.lib.qbx_batch_random_qubo(self._h, len(seeds), states, nsa.ctypes.data_as(_P(ctypes.c_int32)),
                                                int(low), int(high), code, ctypes.byref(o), out)
        if rc not in (QBX_OK, QBX_ERANGE):
            _check(self.lib, rc)
        return out

    def random_qubo_read(self, seeds, ns, low=-100, high=100, dtype=np.float32) -> list:
        """The matrices random_qubo_batch would solve, drawn on the device and copied back
        (int32 for integer dtypes, float64 for float dtypes)."""
        states, nsa, code = self._gen_args(seeds, ns, dtype)
        et = np.int32 if code > 0 else np.float64
        buf = np.empty(int((nsa.astype(np.int64) ** 2).sum()), dtype=et)
        with self.lock:
            rc = self.lib.qbx_random_qubo_read(self._h, len(seeds), states, nsa.ctypes.data_as(_P(ctypes.c_int32)),
                                               int(low), int(high), code, buf.ctypes.data_as(ctypes.c_void_p))
        _check(self.lib, rc)
        out, o = [], 0
        for n in nsa:
            out.append(buf[o:o + n * n].reshape(n, n))
            o += n * n
        return out

    def _tsp_args(self, seeds, n_cities, dtype):
        key = np.dtype(dtype).name
        if key not in ("float32", "float64"):
            raise TypeError(f"TSP cities of dtype {key} cannot be drawn on the device")
        states, nca, code = self._gen_args(seeds, n_cities, dtype)
        return states, nca, code

    def random_tsp_batch(self, seeds, n_cities, low=0, high=100, dtype=np.float32, **kw):
        """Solve TSP.init_random_instance(seed, n_cities, low, high, dtype) instances whose
        QUBOs are built on the device (bit for bit the reference's); returns a sequence of
        Result (as batch())."""
        states, nca, code = self._tsp_args(seeds, n_cities, dtype)
        out = (Result * len(seeds))()
        o = _opts(**kw)
        with self.lock:
            rc = self.lib.qbx_batch_random_tsp(self._h, len(seeds), states, nca.ctypes.data_as(_P(ctypes.c_int32)),
                                               float(low), float(high), code, ctypes.byref(o), out)
        if rc not in (QBX_OK, QBX_ERANGE):
            _check(self.lib, rc)
        return out

    def random_tsp_read(self, seeds, n_cities, low=0, high=100, dtype=np.float32) -> list:
        """The QUBOs random_tsp_batch would solve, built on the device and copied back
        (float64, n_cities^2 square)."""
        states, nca, code = self._tsp_args(seeds, n_cities, dtype)
        ns = nca.astype(np.int64) ** 2
        buf = np.empty(int((ns ** 2).sum()), dtype=np.float64)
        with self.lock:
            rc = self.lib.qbx_random_tsp_read(self._h, len(seeds), states, nca.ctypes.data_as(_P(ctypes.c_int32)),
                                              float(low), float(high), code, buf.ctypes.data_as(_P(ctypes.c_double)))
        _check(self.lib, rc)
        out, o = [], 0
        for n in ns:
            out.append(buf[o:o + n * n].reshape(n, n))
            o += n * n
        return out

    def batch_plan(self, ns, kind: str = "i32", **kw) -> "BatchPlan":
        """Capture the launches of one batch shape in a HIP graph (see BatchPlan)."""
        return BatchPlan(self, ns, kind, **kw)


class BatchPlan:
    """A batch shape (problem sizes, kind, tiling) with its own device memory and a
    captured HIP graph; run() copies new matrices of that shape in and replays the graph.
    Uses the context's stream: one host thread at a time, like the context."""

    def __init__(self, ctx: Context, ns, kind: str = "i32", **kw):
        self.ctx = ctx
        self.kind = kind
        self.ns = np.ascontiguousarray(np.asarray(ns, dtype=np.int32))
        self.dtype = np.int32 if kind == "i32" else np.float64
        self._h = ctypes.c_void_p()
        o = _opts(**kw)
        with ctx.lock:
            rc = ctx.lib.qbx_batch_plan_create(ctx._h, KINDS[kind], int(self.ns.size),
                                               self.ns.ctypes.data_as(_P(ctypes.c_int32)), ctypes.byref(o),
                                               ctypes.byref(self._h))
        _check(ctx.lib, rc)

    def run(self, Qs=None, packed=None):
        """Solve matrices of the planned sizes (a list, or Context.pack output); returns a
        sequence of Result (as Context.batch)."""
        if packed is None:
            packed = Context.pack(Qs, self.dtype)
        data, offsets, ns = packed
        if not np.array_equal(np.asarray(ns, dtype=np.int32), self.ns):
            raise ValueError("matrix sizes differ from the planned batch shape")
        data = np.ascontiguousarray(data, dtype=self.dtype)
        offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        out = (Result * int(self.ns.size))()
        fn, ptr = ((self.ctx.lib.qbx_batch_plan_run_i32, ctypes.c_int32) if self.kind == "i32"
                   else (self.ctx.lib.qbx_batch_plan_run_f64, ctypes.c_double))
        with self.ctx.lock:
            rc = fn(self._h, data.ctypes.data_as(_P(ptr)), offsets.ctypes.data_as(_P(ctypes.c_int64)), out)
        if rc not in (QBX_OK, QBX_ERANGE):
            _check(self.ctx.lib, rc)
        return out

    def close(self):
        if self._h:
            with self.ctx.lock:
                self.ctx.lib.qbx_batch_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_contexts: dict[int, Context] = {}
_ctx_lock = threading.Lock()


def get_context(device: int = 0) -> Context:
    """Process-wide context per HIP device (opened on first use)."""
    with _ctx_lock:
        c = _contexts.get(device)
        if c is None:
            c = Context(device)
            _contexts[device] = c
        return c


# ------------------------------------------------------------------ device='cpu' backend
_CPU_SIGNATURES = {
    "qbx_cpu_solve_i32": (ctypes.c_int, [_P(ctypes.c_int32), ctypes.c_int, _P(Opts), ctypes.c_int, _P(Result)]),
    "qbx_cpu_solve_f64": (ctypes.c_int, [_P(ctypes.c_double), ctypes.c_int, _P(Opts), ctypes.c_int, _P(Result)]),
    "qbx_cpu_solve_minima_i32": (ctypes.c_int, [_P(ctypes.c_int32), ctypes.c_int, _P(Opts), ctypes.c_int,
                                                _P(ctypes.c_int64), _P(ctypes.c_double), _P(Result)]),
    "qbx_cpu_solve_minima_f64": (ctypes.c_int, [_P(ctypes.c_double), ctypes.c_int, _P(Opts), ctypes.c_int,
                                                _P(ctypes.c_double), _P(ctypes.c_double), _P(Result)]),
    "qbx_cpu_last_error": (ctypes.c_char_p, []),
}
CPU_EXPORTED_SYMBOLS = tuple(_CPU_SIGNATURES)
_cpu_lib = None


def load_cpu_library(path: Path | str | None = None) -> ctypes.CDLL:
    """Load libqbx_cpu.so (built in-tree by __graft_entry__.build()).  Raises if it is missing."""
    global _cpu_lib
    with _lib_lock:
        if _cpu_lib is not None and path is None:
            return _cpu_lib
        p = Path(path) if path is not None else CPU_LIB_PATH
        if not p.exists():
            raise NativeError(QBX_EINVAL, f"{p} not found: build it with __graft_entry__.build()")
        lib = ctypes.CDLL(str(p))
        for name, (res, args) in _CPU_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _cpu_lib = lib
        return lib


def _check_cpu(lib, rc: int):
    if rc == QBX_OK:
        return
    msg = lib.qbx_cpu_last_error().decode(errors="replace")
    if rc == QBX_EINVAL:
        raise ValueError(msg)
    if rc == QBX_ERANGE:
        raise OverflowError(msg)
    raise NativeError(rc, msg)


class CpuBackend:
    """device='cpu': libqbx_cpu.so runs the GPU kernels' lane walk on the host, OpenMP
    threads over lanes (threads = 0: OpenMP's default).  Same answers as the GPU; the
    same solve_i32 / solve_f64 / batch calls as Context."""

    device = "cpu"

    def __init__(self, threads: int = 0):
        self.lib = load_cpu_library()
        self.threads = int(threads)
        self.lock = threading.Lock()
        self.tuning_file = None

    def _solve(self, fn, Q, dtype, ptr, kw) -> Result:
        Q = np.ascontiguousarray(Q, dtype=dtype)
        out = Result()
        o = _opts(**kw)
        with self.lock:
            rc = fn(Q.ctypes.data_as(_P(ptr)), Q.shape[0], ctypes.byref(o), self.threads, ctypes.byref(out))
        _check_cpu(self.lib, rc)
        return out

    def solve_i32(self, Q: np.ndarray, **kw) -> Result:
        return self._solve(self.lib.qbx_cpu_solve_i32, Q, np.int32, ctypes.c_int32, kw)

    def solve_f64(self, Q: np.ndarray, **kw) -> Result:
        return self._solve(self.lib.qbx_cpu_solve_f64, Q, np.float64, ctypes.c_double, kw)

    def solve_minima(self, Q: np.ndarray, kind: str = "i32", **kw):
        """As Context.solve_minima, on the host (wave_ms: thread time per prefix)."""
        dtype, fn, ptr = ((np.int32, self.lib.qbx_cpu_solve_minima_i32, ctypes.c_int32) if kind == "i32"
                          else (np.float64, self.lib.qbx_cpu_solve_minima_f64, ctypes.c_double))
        Q = np.ascontiguousarray(Q, dtype=dtype)
        minima, wave_ms, o = _minima_buffers(Q.shape[0], kind, kw)
        out = Result()
        with self.lock:
            rc = fn(Q.ctypes.data_as(_P(ptr)), Q.shape[0], ctypes.byref(o), self.threads,
                    minima.ctypes.data_as(_P(ctypes.c_int64 if kind == "i32" else ctypes.c_double)),
                    wave_ms.ctypes.data_as(_P(ctypes.c_double)), ctypes.byref(out))
        _check_cpu(self.lib, rc)
        return out, minima, wave_ms

    def batch(self, Qs, kind: str = "i32", packed=None, **kw):
        """One solve after another (each solve uses every thread); a problem outside its
        path's range gets status QBX_ERANGE, as in Context.batch."""
        if packed is not None:
            data, offsets, ns = packed
            Qs = [data[o:o + n * n].reshape(n, n) for o, n in zip(offsets, ns)]
        fn = self.solve_i32 if kind == "i32" else self.solve_f64
        out = []
        for Q in Qs:
            try:
                out.append(fn(Q, **kw))
            except OverflowError:
                r = Result()
                r.n, r.status = np.asarray(Q).shape[0], QBX_ERANGE
                out.append(r)
        return out


_cpu_backend = None


def parse_device(device) -> int | str:
    """'cpu' -> 'cpu'; 'cuda' -> 0; 'cuda:N' / N -> N (a HIP device).  ValueError otherwise."""
    if isinstance(device, (int, np.integer)) and not isinstance(device, bool):
        return int(device)
    if isinstance(device, str):
        if device == "cpu":
            return "cpu"
        if device[:4] == "cuda":
            if len(device) == 4:
                return 0
            if device[4] == ":" and device[5:].isdigit():
                return int(device[5:])
    raise ValueError(f"Unknown device name: {device}")


def get_backend(device=0):
    """The solver backend of a device: CpuBackend for 'cpu', else the HIP device's Context."""
    global _cpu_backend
    dev = parse_device(device)
    if dev == "cpu":
        with _ctx_lock:
            if _cpu_backend is None:
                _cpu_backend = CpuBackend()
            return _cpu_backend
    return get_context(dev)
