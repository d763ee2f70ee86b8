"""The kernels' lane walk (csrc/qbx_walk.h), compiled for the host, against the oracle.

libqbx_emu.so runs the same walk_lane_* templates the gfx950 kernels run, lane by lane,
on the same bit-ordered tables.  Every (n, walk bits L, tie width B, inner bits M) shape
is checked here on the CPU, so lane prefixes, block keys, tie ranks and outer transitions
are pinned before the GPU runs them.
"""
import ctypes
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import KINDS, PAIR_BOUND, energy_bound, itertools_min, pair_case, random_instance

EMU = Path(__file__).resolve().parent / "_build" / "libqbx_emu.so"
P = ctypes.POINTER


@pytest.fixture(scope="module")
def emu():
    if not EMU.exists():
        import build_native
        build_native.build_emu()
    L = ctypes.CDLL(str(EMU))
    L.qbx_emu_solve_i32.argtypes = [P(ctypes.c_int32), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    P(ctypes.c_int64), P(ctypes.c_uint64), P(ctypes.c_uint64)]
    L.qbx_emu_solve_f64.argtypes = [P(ctypes.c_double), ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    P(ctypes.c_double), P(ctypes.c_uint64), P(ctypes.c_uint64)]
    L.qbx_emu_take_stats.argtypes = [P(ctypes.c_uint64), P(ctypes.c_uint64)]
    L.qbx_emu_state_to_rank.restype = ctypes.c_uint64
    L.qbx_emu_state_to_rank.argtypes = [ctypes.c_uint64, ctypes.c_int]
    L.qbx_emu_rank_to_state.restype = ctypes.c_uint64
    L.qbx_emu_rank_to_state.argtypes = [ctypes.c_uint64, ctypes.c_int]
    return L


def run_i32(emu, Q, B, L, M):
    Q = np.ascontiguousarray(Q, dtype=np.int32)
    E, r, s = ctypes.c_int64(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = emu.qbx_emu_solve_i32(Q.ctypes.data_as(P(ctypes.c_int32)), Q.shape[0], B, L, M,
                               ctypes.byref(E), ctypes.byref(r), ctypes.byref(s))
    assert rc == 0, rc
    return E.value, s.value


def run_f64(emu, Q, B, L, M):
    Q = np.ascontiguousarray(Q, dtype=np.float64)
    E, r, s = ctypes.c_double(), ctypes.c_uint64(), ctypes.c_uint64()
    rc = emu.qbx_emu_solve_f64(Q.ctypes.data_as(P(ctypes.c_double)), Q.shape[0], B, L, M,
                               ctypes.byref(E), ctypes.byref(r), ctypes.byref(s))
    assert rc == 0, rc
    return E.value, s.value


def test_rank_helpers_match_oracle(emu):
    for B in (0, 3, 9, 20):
        for s in list(range(512)) + [0xFFFFF, 0x12345678]:
            r = emu.qbx_emu_state_to_rank(s, B)
            assert r == oracle.rank_of_state(s, B)
            assert emu.qbx_emu_rank_to_state(r, B) == s


@pytest.mark.parametrize("M", [1, 4, 6, 7])
def test_every_shape_small(emu, M):
    rng = np.random.default_rng(100 + M)
    for trial in range(40):
        n = int(rng.integers(M, 12))
        Q = random_instance(rng, n, KINDS[trial % len(KINDS)])
        B = int(rng.integers(0, n + 1)) if trial % 2 else min(n, 20)
        ref = itertools_min(Q, B)
        for L in range(M, n + 1):
            assert run_i32(emu, Q, B, L, M) == ref, (n, B, L, M)


@pytest.mark.parametrize("M", [6, 7])
def test_medium_against_brute(emu, M):
    rng = np.random.default_rng(7 + M)
    for trial in range(10):
        n = int(rng.integers(14, 21))
        Q = random_instance(rng, n, KINDS[trial % 4])
        B = [min(n, 20), 0, 3, n, M - 1, M + 2][trial % 6]
        ref = oracle.brute_i64(Q, B)
        for L in sorted({M, M + 3, M + 4, M + 5, min(n, M + 13)}):
            if L <= n:
                assert run_i32(emu, Q, B, L, M) == ref, (n, B, L, M)


def _float_case(rng, n, trial):
    kind = trial % 6
    if kind == 0:      # exact ties: integer-valued float matrix
        return random_instance(rng, n, "maxcut").astype(np.float64)
    if kind == 1:      # one huge entry: the small ones quantize to zero, the filter passes much
        Q = rng.standard_normal((n, n)) * 1e-3
        Q[0, n - 1] = 1e9
        return Q
    if kind == 2:      # tiny magnitudes (scale far above 2^31)
        return rng.standard_normal((n, n)) * 1e-200
    if kind == 3:      # near-ties: two copies of a state's energy within 1e-12
        Q = np.round(rng.standard_normal((n, n)) * 8) / 8
        Q[1, 1] += 1e-12
        return Q
    return rng.standard_normal((n, n)).astype(np.float32).astype(np.float64)


def take_stats(emu):
    f, r = ctypes.c_uint64(), ctypes.c_uint64()
    emu.qbx_emu_take_stats(ctypes.byref(f), ctypes.byref(r))
    return f.value, r.value


@pytest.mark.parametrize("M", [1, 4, 6, 7])
def test_float_filter_walk(emu, M):
    """int32 filter over quantized tables + FP64 re-walk gives the FP64 brute-force answer,
    through every path of the deferred-candidate buffer (overflow flushes included)."""
    rng = np.random.default_rng(31 + M)
    take_stats(emu)
    flushes = rewalks = 0
    for trial in range(12):
        n = int(rng.integers(max(M, 4), 17))
        Q = _float_case(rng, n, trial)
        B = [min(n, 20), 0, n // 2][trial % 3]
        E, s = oracle.brute_f64(Q, B)
        for L in sorted({M, min(n, M + 4), min(n, M + 9)}):
            E2, s2 = run_f64(emu, Q, B, L, M)
            # FP64 sums in another order differ by rounding only: |dE| <= 1e-13 sum|Q|
            assert s2 == s and E2 == pytest.approx(E, rel=0, abs=1e-13 * np.abs(Q).sum() + 1e-300), \
                (n, B, L, M, trial)
            f, r = take_stats(emu)
            flushes, rewalks = flushes + f, rewalks + r
    # a lane with hundreds of blocks that tie after quantization overflows its record buffer
    n = 16
    Q = _float_case(rng, n, 1)
    E, s = oracle.brute_f64(Q, 16)
    E2, s2 = run_f64(emu, Q, 16, min(n, M + 9), M)
    assert s2 == s and E2 == pytest.approx(E, rel=0, abs=1e-13 * np.abs(Q).sum())
    f, r = take_stats(emu)
    assert f > 0 and r > 0
    assert rewalks > 0



def take_qrp(emu):
    lanes, groups = ctypes.c_uint64(), ctypes.c_uint64()
    emu.qbx_emu_take_qrp_stats(ctypes.byref(lanes), ctypes.byref(groups))
    return lanes.value, groups.value


@pytest.mark.parametrize("M", [1, 4, 6, 7])
@pytest.mark.parametrize("slack,shift_delta", [(-1, 0), (0, 0), (-1, -3), (-1, 3), (-1, 6), (40, 0)])
def test_packed_float_filter(emu, M, slack, shift_delta):
    """The packed float filter (coarse 16-bit block-relative energies over the quantized
    tables) gives the int32 filter's answer exactly (the emulation returns -7 otherwise)
    and the FP64 brute force's; with the guard's slack cut (0: every group in int32; 40:
    some) and at finer (-3: more int32 groups) and coarser (+3: more candidates) scales."""
    emu.qbx_emu_set_qrp.argtypes = [ctypes.c_int32, ctypes.c_int32]
    emu.qbx_emu_take_qrp_stats.argtypes = [P(ctypes.c_uint64), P(ctypes.c_uint64)]
    emu.qbx_emu_set_qrp(slack, shift_delta)
    try:
        rng = np.random.default_rng(61 + M)
        take_qrp(emu)
        lanes = groups = 0
        for trial in range(10):
            n = int(rng.integers(M + 6, 17)) if M + 6 < 17 else 16
            Q = _float_case(rng, n, trial) if trial % 2 else rng.standard_normal((n, n))
            B = [min(n, 20), 0, n // 2][trial % 3]
            E, s = oracle.brute_f64(Q, B)
            for L in sorted({M + 2, min(n, M + 5), min(n, M + 6), min(n, M + 10)}):
                E2, s2 = run_f64(emu, Q, B, L, M)
                assert s2 == s and E2 == pytest.approx(E, rel=0, abs=1e-13 * np.abs(Q).sum() + 1e-300), \
                    (n, B, L, M, trial)
                got = take_qrp(emu)
                lanes, groups = lanes + got[0], groups + got[1]
        assert lanes > 0
        if slack == 0:
            assert groups > 0
    finally:
        emu.qbx_emu_set_qrp(-1, 0)


@pytest.mark.parametrize("M", [1, 4, 6, 7])
def test_pair_walk_matches_int32_walk(emu, M):
    """Problems with bound <= 15359 take the packed pair walk (two states per 32-bit word).
    The emulator runs it beside the int32 walk for every lane and fails on any differing
    lane key; the answers must also be the brute force's."""
    rng = np.random.default_rng(500 + M)
    emu.qbx_emu_take_pair_lanes.restype = ctypes.c_uint64
    emu.qbx_emu_take_pair_lanes()
    at_bound = 0
    for trial in range(30):
        n = int(rng.integers(M + 1, 15))
        Q = pair_case(rng, n, trial)
        assert energy_bound(Q) <= PAIR_BOUND
        at_bound += energy_bound(Q) == PAIR_BOUND
        B = [min(n, 20), 0, int(rng.integers(0, n + 1))][trial % 3]
        ref = oracle.brute_i64(Q, B)
        for L in sorted({M + 1, min(n, M + 5), min(n, M + 6), n}):
            assert run_i32(emu, Q, B, L, M) == ref, (n, B, L, M, trial)
            assert emu.qbx_emu_take_pair_lanes() == 2 ** (n - L), (n, L, M)
    assert at_bound >= 10
    # above the bound, or with no outer bit to pair on, the int32 walk runs alone
    Q = np.full((10, 10), 200)
    assert energy_bound(Q) > PAIR_BOUND
    assert run_i32(emu, Q, 10, M + 2, M) == oracle.brute_i64(Q, 10)
    Q = random_instance(rng, 9, "ternary")
    assert run_i32(emu, Q, 9, M, M) == oracle.brute_i64(Q, 9)
    assert emu.qbx_emu_take_pair_lanes() == 0


def take_rp(emu):
    lanes, groups = ctypes.c_uint64(), ctypes.c_uint64()
    emu.qbx_emu_take_rp_stats(ctypes.byref(lanes), ctypes.byref(groups))
    return lanes.value, groups.value


def _rp_case(rng, n, trial):
    """Integer problems beyond the pair bound whose in-block energy differences mostly fit
    16 bits: random int16-range QUBOs of several scales."""
    scale = [127, 300, 600, 1200][trial % 4]
    U = rng.integers(-scale, scale + 1, (n, n))
    return np.triu(U) + np.triu(U, 1).T


@pytest.mark.parametrize("M", [1, 4, 6, 7])
@pytest.mark.parametrize("slack", [-1, 0, 3000])
def test_rebased_pair_walk_matches_int32_walk(emu, M, slack):
    """The rebased pair walk (energies relative to each block's base, a per-group guard on
    the fields, int32 groups where it fails) gives every lane the int32 walk's key.  slack
    -1 is the device's guard; 0 forces every group of a lane whose fields are not all zero
    into int32; 3000 mixes both within lanes."""
    emu.qbx_emu_set_rp_slack.argtypes = [ctypes.c_int32]
    emu.qbx_emu_set_rp_slack(slack)
    take_rp(emu)
    lanes = groups = 0
    try:
        rng = np.random.default_rng(700 + M + slack)
        for trial in range(16):
            n = int(rng.integers(M + 2, 16))
            Q = _rp_case(rng, n, trial) if trial % 4 != 3 else pair_case(rng, n, trial)
            B = [min(n, 20), 0, int(rng.integers(0, n + 1))][trial % 3]
            ref = oracle.brute_i64(Q, B)
            for L in sorted({M + 1, min(n, M + 5), min(n, M + 6), n}):
                assert run_i32(emu, Q, B, L, M) == ref, (n, B, L, M, trial)
                a, g = take_rp(emu)
                lanes, groups = lanes + a, groups + g
    finally:
        emu.qbx_emu_set_rp_slack(-1)
    assert lanes > 0
    if slack >= 0:
        assert groups > 0       # the int32 groups ran


def test_rebased_pair_guard_fails_sometimes(emu):
    """A problem the device sends down the rebased pair walk (its heuristic accepts it) whose
    guard fails in some groups: the int32 groups run under the device's own slack, and the
    answers stay exact."""
    from helpers import guard_failing_case
    rng = np.random.default_rng(77)
    take_rp(emu)
    for n in (17, 18):
        Q = guard_failing_case(rng, n)
        assert energy_bound(Q) > PAIR_BOUND
        ref = oracle.brute_i64(Q, None)
        assert run_i32(emu, Q, min(n, 20), 16, 7) == ref
        lanes, groups = take_rp(emu)
        total = lanes * (2 ** (16 - 7 - 1) // 16)
        assert lanes > 0 and 0 < groups < total, (lanes, groups, total)
