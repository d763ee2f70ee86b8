"""Shared test helpers: small independent checks and instance makers."""
from __future__ import annotations

import itertools

import numpy as np


def invgray(v: int) -> int:
    r = 0
    while v:
        r ^= v
        v >>= 1
    return r


def rank_of(s: int, B: int) -> int:
    return (invgray(s >> B) << B) | (s & ((1 << B) - 1))


def itertools_min(Q, B=None):
    """Pure-Python enumeration of all 2^n states, exact (E, rank) minimum (n <= 12)."""
    Q = np.asarray(Q)
    n = Q.shape[0]
    B = min(n, 20) if B is None else B
    best = None
    for bits in itertools.product((0, 1), repeat=n):
        x = np.array(bits)
        E = x @ Q @ x
        E = int(E) if np.issubdtype(Q.dtype, np.integer) else float(E)
        s = int("".join(map(str, bits)), 2) if n else 0
        key = (E, rank_of(s, B))
        if best is None or key < best[0]:
            best = (key, s)
    return best[0][0], best[1]


def random_instance(rng, n, kind):
    if kind == "uniform":
        return rng.integers(-127, 128, (n, n))
    if kind == "maxcut":
        Q = np.zeros((n, n), dtype=np.int64)
        for u, v in zip(*np.nonzero(np.triu(rng.random((n, n)) < 0.5, 1))):
            Q[u, u] -= 1
            Q[v, v] -= 1
            Q[u, v] += 2
        return Q
    if kind == "ternary":
        return rng.integers(-1, 2, (n, n))
    if kind == "sparse":
        return rng.integers(-5, 6, (n, n)) * (rng.random((n, n)) < 0.3)
    if kind == "zero":
        return np.zeros((n, n), dtype=np.int64)
    if kind == "positive":
        return rng.integers(1, 10, (n, n))
    raise ValueError(kind)


KINDS = ("uniform", "maxcut", "ternary", "sparse", "zero", "positive")


PAIR_BOUND = 15359


def energy_bound(Q):
    Q = np.asarray(Q, dtype=np.int64)
    up = np.triu(Q, 1) + np.triu(Q.T, 1)
    return int(np.abs(np.diag(Q)).sum() + np.abs(up).sum())


def pair_case(rng, n, trial):
    """Integer problems whose energies fit the packed pair walk, including both ends of its
    range (an all-ones state reaching -PAIR_BOUND or +PAIR_BOUND)."""
    kind = trial % 5
    if kind == 0:
        return random_instance(rng, n, "maxcut")
    if kind == 1:
        return random_instance(rng, n, "ternary")
    if kind == 2:      # every entry <= 0: the all-ones state sits at -bound
        Q = -rng.integers(0, 60, (n, n))
    elif kind == 3:    # every entry >= 0: the all-ones state sits at +bound
        Q = rng.integers(0, 60, (n, n))
    else:
        Q = rng.integers(-90, 91, (n, n))
    Q = np.asarray(Q, dtype=np.int64)
    b = energy_bound(Q)
    if b > PAIR_BOUND // 2:     # scale into range
        Q = Q * (PAIR_BOUND // 2) // b
    # grow |Q[0, 0]| until the bound is exactly PAIR_BOUND
    gap = PAIR_BOUND - energy_bound(Q)
    Q[0, 0] += -gap if Q[0, 0] < 0 or kind == 2 else gap
    return Q


def guard_failing_case(rng, n, w=600, noise=3):
    """An integer problem beyond the pair bound that takes the rebased pair walk and whose
    guard fails in 1/16 of the groups: every inner field (state bits 0..6, variables
    n-1..n-7) is w times the number of set bits among state bits 12..15, so the fields' sum
    reaches 28 w when all four are set, beyond the guard's room; small noise elsewhere."""
    Q = rng.integers(-noise, noise + 1, (n, n))
    Q = np.triu(Q) + np.triu(Q, 1).T
    for a in range(12, 16):
        for k in range(7):
            i, j = n - 1 - a, n - 1 - k
            Q[i, j] = Q[j, i] = w // 2     # W[a][k] = Q_ij + Q_ji = w (w even)
    return Q
