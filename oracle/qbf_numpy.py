"""numpy restatement of the reference QBF algorithm -- TEST INFRASTRUCTURE ONLY.

Follows src/qubo_lib/solvers/qbf.py step by step on the CPU:
  define_chunk_size          B = min(N, 20), A = N - B                qbf.py:429-446
  split_matrix_to_blocks     Q_0A, Q_0B, Q_1A, Q_1B                   qbf.py:422-427
  build_energy_vector        recursive energy table of Q_1B           qbf.py:516-548
  fill_all_state_matrix      X_B, MSB first                           qbf.py:373-384
  compute_ABBA_chunked       (Q_0B + Q_1A^T) @ X_B^T in 16 chunks     qbf.py:264-289
  gray_code_iterator         g = i ^ (i >> 1), flipped bit, sign      qbf.py:291-325
  iterate_prefix_interval    chunk +-= ABBA row, argmin, update       qbf.py:188-230
  update_global_minima       prefix_E, strict '<', reset at value 0   qbf.py:395-412
Integer Q accumulates in int64 (exact; the reference uses Q.dtype and may wrap), float Q
in the input dtype until the first correction and float64 after it, as the reference
does (qbf.py:240-241).  It is the CPU baseline timed by bench.py and a second restatement
checked against oracle/qbf_oracle.c in the tests.
"""
from __future__ import annotations

import time

import numpy as np


def fill_all_state_matrix(mx: np.ndarray) -> None:
    n = mx.shape[1]
    assert mx.shape == (2 ** n, n)
    if n == 0:
        return
    mx[1, -1] = 1
    for i in range(1, n):
        col = n - i
        row = 2 ** i
        mx[row:2 * row, col:] = mx[:row, col:]
        mx[row:2 * row, col - 1] = 1


def gray_code_iterator(bits_count: int):
    """Yields (flipped_bit_idx, sign, value, bit_array) (qbf.py:291-325)."""
    bit_array = np.zeros(bits_count)
    prev = 0
    yield 0, 0, 0, bit_array.copy()
    for i in range(1, 1 << bits_count):
        g = i ^ (i >> 1)
        sign = 1 if g > prev else -1
        flipped = (g ^ prev).bit_length() - 1
        bit_array[bits_count - (flipped + 1)] += sign
        prev = g
        yield flipped, sign, g, bit_array.copy()


def _energies_direct(Q: np.ndarray) -> np.ndarray:
    n = Q.shape[0]
    X = np.zeros((2 ** n, n), dtype=Q.dtype)
    fill_all_state_matrix(X)
    return np.einsum("ij,jk,ik->i", X, Q, X)


def build_energy_vector(Q: np.ndarray, A: int, B: int) -> np.ndarray:
    """Energy of every state of Q, indexed by the state integer (qbf.py:516-548)."""
    N = Q.shape[0]
    out = np.zeros(2 ** N, dtype=Q.dtype)
    for chunk, value in _energy_chunks(Q, A, B):
        out[value * 2 ** B:(value + 1) * 2 ** B] = chunk
    return out


def _energy_chunks(Q, A, B):
    Q_0A = Q[:A, :A]
    Q_0B = Q[:A, A:]
    Q_1A = Q[A:, :A]
    Q_1B = Q[A:, A:]
    X_B = np.zeros((2 ** B, B), dtype=Q.dtype)
    fill_all_state_matrix(X_B)
    ABBA = (Q_0B + Q_1A.T) @ X_B.T
    chunk = _energies_direct(Q_1B) if B <= 10 else build_energy_vector(Q_1B, B - 10, 10)
    for flipped, sign, value, bits in gray_code_iterator(A):
        row = A - flipped - 1
        if sign == 1:
            chunk = chunk + ABBA[row]
        elif sign == -1:
            chunk = chunk - ABBA[row]
        prefix_E = bits @ Q_0A @ bits.T
        yield prefix_E + chunk, value


class QBFNumpy:
    """One QBF solve; `steps` limits the prefix walk to its first `steps` prefixes (a
    bounded sample of the full 2^A walk, for timing)."""

    def __init__(self, Q, B=None, correction_period=1024):
        Q = np.asarray(Q)
        if np.issubdtype(Q.dtype, np.integer):
            Q = Q.astype(np.int64)
        self.Q = Q
        self.N = Q.shape[0]
        self.B = min(self.N, 20) if B is None else int(B)
        self.A = self.N - self.B
        self.correction_period = correction_period
        self.build_seconds = None

    def build_cache(self):
        t0 = time.perf_counter()
        A, B, Q = self.A, self.B, self.Q
        self.Q_0A = Q[:A, :A]
        Q_0B, Q_1A, Q_1B = Q[:A, A:], Q[A:, :A], Q[A:, A:]
        tmp_A = 5 if B > 10 else 0
        self.diag_BB = build_energy_vector(Q_1B, tmp_A, B - tmp_A)
        X_B = np.zeros((2 ** B, B), dtype=Q.dtype)
        fill_all_state_matrix(X_B)
        W = Q_0B + Q_1A.T
        self.ABBA = np.empty((A, 2 ** B), dtype=Q.dtype)
        cnt = 16 if B >= 4 else 1
        lil = (2 ** B) // cnt
        for i in range(cnt):
            self.ABBA[:, i * lil:(i + 1) * lil] = W @ X_B[i * lil:(i + 1) * lil].T
        self.build_seconds = time.perf_counter() - t0

    def run(self, steps=None):
        """Returns (glob_E, state, prefix_steps_done, seconds_in_prefix_walk)."""
        if not hasattr(self, "ABBA"):
            self.build_cache()
        A, B = self.A, self.B
        chunk = self.diag_BB.copy()
        glob_E, glob_pref, glob_X = 0, 0, 0
        t0 = time.perf_counter()
        done = 0
        for i, (flipped, sign, value, bits) in enumerate(gray_code_iterator(A)):
            if steps is not None and i >= steps:
                break
            row = A - flipped - 1
            if sign == 1:
                chunk += self.ABBA[row]
            elif sign == -1:
                chunk -= self.ABBA[row]
            if (i + 1) % self.correction_period == 0:
                chunk = self.diag_BB.copy() + np.array(bits) @ self.ABBA
                if np.issubdtype(self.Q.dtype, np.integer):
                    chunk = np.rint(chunk).astype(np.int64)
            x = int(chunk.argmin())
            best_chunk_E = chunk[x]
            if value == 0:
                glob_E, glob_pref, glob_X = 0, 0, 0
            prefix_E = bits @ self.Q_0A @ bits.T
            if best_chunk_E + prefix_E < glob_E:
                glob_E, glob_pref, glob_X = best_chunk_E + prefix_E, value, x
            done += 1
        seconds = time.perf_counter() - t0
        return glob_E, (glob_pref << B) | glob_X, done, seconds


def solve(Q, B=None, correction_period=1024):
    """Full QBF solve: returns (E, state)."""
    E, s, _, _ = QBFNumpy(Q, B, correction_period).run()
    if np.issubdtype(np.asarray(Q).dtype, np.integer):
        E = int(round(float(E)))
    else:
        E = float(E)
    return E, s
