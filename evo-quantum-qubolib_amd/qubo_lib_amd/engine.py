"""Host-side dispatch from a dense Q to the gfx950 kernels (libqbx.so).

Arithmetic path per problem:
  * integer Q with energy bound  sum|Q_ii| + sum_{i<j}|Q_ij + Q_ji|  <= 2^31 - 2:
      exact int32 walk (qbx_solve_i32).  Q is passed in folded upper-triangular form
      (Q'_ij = Q_ij + Q_ji, i < j), which has the same energies and fits int32 whenever
      the bound does.
  * integer Q with bound < 2^53: the FP64 walk, exact on integers (qbx_solve_f64).
  * float Q whose entries are all integers (MaxCut's default float32 matrices): the
      exact int32 walk, energies returned as floats.
  * float Q whose entries are all multiples of one power of two u = 2^-k (weights / 4,
      quarter- or half-integer matrices), with an energy bound of at most 2^31 - 2 units
      of u: the same, on Q / u (an exact integer problem), energies u * E_int -- exact
      values, and the same order and ties as the float path gives them.
  * other float Q: the FP64 walk on the float64 values of Q (the int32 walk over tables
      quantized to 2^s units filters blocks with a proven error bound; every block it
      passes is re-walked in FP64).
  * anything else: OverflowError.

`device` is a HIP device (an int, 'cuda' or 'cuda:N': libqbx.so, never a CPU fallback) or
'cpu' (libqbx_cpu.so: the same lane walk on the host, same answers).

Energies of integer problems come back as Python ints, of float problems as floats.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _native

MAX_N = 56
INT32_BOUND = (1 << 31) - 2
EXACT_F64_BOUND = 1 << 53


@dataclass
class Prepared:
    """A problem matrix validated and converted for one kernel path."""
    n: int
    kind: str                 # 'i32' | 'f64'
    integer: bool             # the problem's values are integers
    matrix: np.ndarray        # int32 (folded) or float64
    bound: float
    unit: float = 1.0         # i32 kind: energies are unit * the integer walk's (a power of two)


@dataclass
class RawResult:
    E: object                 # int for integer problems, float otherwise
    state: int
    rank: int
    n: int
    tie_B: int
    kind: str
    info: dict = field(default_factory=dict)


def default_tie_B(n: int) -> int:
    """The reference CPU QBF's suffix width B = min(N, 20) (qbf.py:443-446)."""
    return min(n, 20)


def prepare(Q, lattice: bool = True) -> Prepared:
    """Validate Q and choose its arithmetic path.  lattice=False keeps float matrices whose
    entries are multiples of a power of two on the float path (same answer, slower)."""
    Q = np.asarray(Q)
    if Q.ndim != 2 or Q.shape[0] != Q.shape[1]:
        raise ValueError(f"Q must be a square matrix, got shape {Q.shape}")
    n = Q.shape[0]
    if n < 1 or n > MAX_N:
        raise ValueError(f"exhaustive solve needs 1 <= N <= {MAX_N}, got N={n}")
    if Q.dtype == np.bool_:
        Q = Q.astype(np.int64)
    if np.issubdtype(Q.dtype, np.integer):
        if Q.dtype.itemsize <= 4 or (-2 ** 31 <= int(Q.min()) and int(Q.max()) <= 2 ** 31):
            # entries within +-2^31: every sum below fits int64 (n <= 56), vectorised
            Qi = Q.astype(np.int64)
            up = np.triu(Qi, 1) + np.triu(Qi.T, 1)      # Q_ij + Q_ji for i < j, exact
            diag = np.diag(Qi)
            bound = int(np.abs(diag).sum()) + int(np.abs(up).sum())
        else:
            Qo = Q.astype(object)                       # arbitrary int64 / uint64: exact in Python ints
            up = np.triu(Qo, 1) + np.triu(Qo.T, 1)
            diag = np.diag(Qo)
            bound = int(sum(abs(v) for v in diag)) + int(sum(abs(v) for v in up.ravel()))
        folded = up + np.diag(diag)
        if bound <= INT32_BOUND:
            return Prepared(n, "i32", True, np.ascontiguousarray(folded.astype(np.int64).astype(np.int32)), bound)
        if bound < EXACT_F64_BOUND:
            return Prepared(n, "f64", True, np.ascontiguousarray(folded.astype(np.float64)), bound)
        raise OverflowError(f"integer Q with energy bound {bound} >= 2^53 cannot be solved exactly")
    if np.issubdtype(Q.dtype, np.floating):
        Qf = Q.astype(np.float64)
        if not np.all(np.isfinite(Qf)):
            raise ValueError("Q contains non-finite values")
        if np.all(Qf == np.rint(Qf)) and np.abs(Qf).max(initial=0.0) < 2.0 ** 53:
            # integer-valued float Q (e.g. MaxCut's default float32 matrices): the energies
            # are integers, so the exact int32 walk gives the same answer; E stays a float.
            # Integer bounds beyond the exact integer paths fall through to the float path,
            # which solves any finite Q.
            try:
                p = prepare(Qf.astype(np.int64))
            except OverflowError:
                p = None
            if p is not None and p.kind == "i32":
                return Prepared(p.n, "i32", False, p.matrix, p.bound)
        k = lattice_bits(Qf) if lattice else -1
        if 0 < k <= 62:
            # Q = u * Z with u = 2^-k and Z integral: the integer walks on Z, exactly
            Z = np.ldexp(Qf, k)
            if np.abs(Z).max(initial=0.0) < 2.0 ** 53:
                try:
                    p = prepare(Z.astype(np.int64))
                except OverflowError:
                    p = None
                if p is not None and p.kind == "i32":
                    return Prepared(p.n, "i32", False, p.matrix, p.bound, unit=float(np.ldexp(1.0, -k)))
        bound = float(np.abs(np.diag(Qf)).sum() + np.abs(np.triu(Qf, 1) + np.triu(Qf.T, 1)).sum())
        return Prepared(n, "f64", False, np.ascontiguousarray(Qf), bound)
    raise TypeError(f"unsupported Q dtype {Q.dtype}")


def lattice_bits(Qf: np.ndarray) -> int:
    """Smallest k >= 0 with Qf * 2^k all integers (every finite double is a multiple of a
    power of two: its odd integer mantissa times 2^e), or -1 when it would exceed 1074."""
    v = Qf[Qf != 0.0]
    if v.size == 0:
        return 0
    m, e = np.frexp(np.abs(v))                  # |v| = m 2^e, 0.5 <= m < 1
    M = np.ldexp(m, 53).astype(np.int64)        # the 53-bit integer mantissa, exact
    tz = np.log2((M & -M).astype(np.float64)).astype(np.int64)   # its trailing zero bits
    k = int((53 - e - tz).max())                # |v| = (M >> tz) 2^(e - 53 + tz)
    return max(k, 0) if k <= 1074 else -1


def energy_of(p: Prepared, E_int) -> float:
    """Energy of an integer-walk result of a float problem (unit * E_int, exact)."""
    return float(E_int) * p.unit


def _finish(p: Prepared, r: _native.Result) -> RawResult:
    if p.integer:
        E = int(r.E_int) if p.kind == "i32" else int(round(r.E))
    elif p.kind == "i32":
        E = energy_of(p, r.E_int)
    else:
        E = float(r.E)
    info = dict(kernel_ms=r.kernel_ms, states=int(r.states), tiles=int(r.tiles), walk_bits=r.walk_bits,
                inner_bits=r.inner_bits, lane_bits=r.lane_bits, path=p.kind, bound=p.bound,
                walk=_native.PATHS.get(r.path))
    return RawResult(E=E, state=int(r.state), rank=int(r.rank), n=p.n, tie_B=r.tie_B, kind=p.kind, info=info)


def solve_matrix(Q, device=0, tie_B=None, walk_bits=None, inner_bits=None, float_filter=None) -> RawResult:
    """Exact minimum of x^T Q x and its argmin under the reference's tie rule.
    float_filter (float problems): None (by size), "packed" or "int32"; the answer is the same."""
    p = Q if isinstance(Q, Prepared) else prepare(Q)
    if tie_B is not None and not (0 <= int(tie_B) <= p.n):
        raise ValueError(f"B={tie_B} outside [0, N={p.n}]")
    ctx = _native.get_backend(device)
    kw = dict(tie_B=tie_B, walk_bits=walk_bits, inner_bits=inner_bits, float_filter=float_filter)
    r = ctx.solve_i32(p.matrix, **kw) if p.kind == "i32" else ctx.solve_f64(p.matrix, **kw)
    return _finish(p, r)


def solve_matrix_minima(Q, device=0, tie_B=None, walk_bits=None, inner_bits=None, float_filter=None):
    """solve_matrix's answer and, for the reference's per-prefix metrics (qbf.py:342-351,
    362-365, 409-410), per reference prefix (top N - B bits) in Gray visiting order i
    (prefix i ^ (i >> 1)): its least energy (int for integer problems' values) and the
    device time spent on it in ms.  Returns (RawResult, minima, wave_ms); needs B >= 1
    and N - B <= 28.  Lanes are capped at B bits (walk_bits / inner_bits too)."""
    p = Q if isinstance(Q, Prepared) else prepare(Q)
    if tie_B is not None and not (0 <= int(tie_B) <= p.n):
        raise ValueError(f"B={tie_B} outside [0, N={p.n}]")
    ctx = _native.get_backend(device)
    r, minima, wave_ms = ctx.solve_minima(p.matrix, kind=p.kind, tie_B=tie_B, walk_bits=walk_bits,
                                          inner_bits=inner_bits, float_filter=float_filter)
    if p.kind == "f64" and p.integer:
        minima = np.rint(minima).astype(np.int64)
    elif p.kind == "i32" and not p.integer:
        minima = minima.astype(np.float64) * p.unit
    return _finish(p, r), minima, wave_ms


def solve_part(Q, part: int, parts: int, device=0, tie_B=None, walk_bits=None,
               inner_bits=None, float_filter=None) -> RawResult:
    """The (E, rank) minimum over one of `parts` equal shares of the solve's state space (one
    GPU's share of a solve spread over `parts` GPUs); merge_parts of all shares is
    solve_matrix's answer.  The share boundaries follow the solve's tiling, so every share
    must be solved with the same walk_bits / inner_bits."""
    p = Q if isinstance(Q, Prepared) else prepare(Q)
    if tie_B is not None and not (0 <= int(tie_B) <= p.n):
        raise ValueError(f"B={tie_B} outside [0, N={p.n}]")
    if _native.parse_device(device) == "cpu":
        raise ValueError("solve_part is one GPU's share of a solve: it needs a 'cuda' device")
    ctx = _native.get_backend(device)
    r = ctx.solve_part(p.matrix, part, parts, kind=p.kind, tie_B=tie_B, walk_bits=walk_bits,
                       inner_bits=inner_bits, float_filter=float_filter)
    return _finish(p, r)


def merge_parts(parts) -> RawResult:
    """Lexicographic minimum of (E, rank) over share results: the tie rule of the whole solve
    (an equal energy goes to the lower rank, whichever share found it)."""
    parts = list(parts)
    if not parts:
        raise ValueError("no parts to merge")
    best = min(parts, key=lambda r: (r.E, r.rank))
    info = dict(best.info, states=sum(r.info.get("states", 0) for r in parts),
                tiles=sum(r.info.get("tiles", 0) for r in parts), parts=len(parts),
                kernel_ms=max(r.info.get("kernel_ms", 0.0) for r in parts))
    return RawResult(E=best.E, state=best.state, rank=best.rank, n=best.n, tie_B=best.tie_B, kind=best.kind,
                     info=info)


def raw_from_native(r: _native.Result, integer: bool) -> RawResult:
    """RawResult of a problem whose matrix never reached the host (device-drawn batches):
    integer problems on the int32 path, float problems on the FP64 path."""
    if r.status != 0:
        raise OverflowError(f"N={r.n}: energy bound outside the {'int32' if integer else 'float'} path")
    kind = "i32" if integer else "f64"
    E = int(r.E_int) if integer else float(r.E)
    info = dict(kernel_ms=r.kernel_ms, states=int(r.states), tiles=int(r.tiles), walk_bits=r.walk_bits,
                inner_bits=r.inner_bits, lane_bits=r.lane_bits, path=kind, bound=None,
                walk=_native.PATHS.get(r.path))
    return RawResult(E=E, state=int(r.state), rank=int(r.rank), n=int(r.n), tie_B=r.tie_B, kind=kind, info=info)


def solve_matrices(Qs, device=0, tie_B=None, walk_bits=None, inner_bits=None, float_filter=None):
    """Many problems, one persistent launch per arithmetic path (on 'cpu': one solve after
    another).  Returns RawResults in input order.  `tie_B` may be a single value
    (None = min(N, 20) per problem)."""
    preps = [q if isinstance(q, Prepared) else prepare(q) for q in Qs]
    out = [None] * len(preps)
    ctx = _native.get_backend(device)
    for kind in ("i32", "f64"):
        idx = [i for i, p in enumerate(preps) if p.kind == kind]
        if not idx:
            continue
        res = ctx.batch([preps[i].matrix for i in idx], kind=kind, tie_B=tie_B, walk_bits=walk_bits,
                        inner_bits=inner_bits, float_filter=float_filter)
        for i, r in zip(idx, res):
            if r.status != 0:
                raise OverflowError(f"problem {i}: energy bound outside the {kind} path")
            out[i] = _finish(preps[i], r)
    return out


def state_bits(state: int, n: int) -> np.ndarray:
    """State integer -> int32 array of n bits, x[0] first (qbf.py:387-393)."""
    return np.array([(state >> (n - 1 - i)) & 1 for i in range(n)], dtype=np.int32)
