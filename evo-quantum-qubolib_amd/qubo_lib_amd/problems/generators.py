"""Synthetic workloads used by the benchmarks and tests (not part of the reference API).

  symmetric_int_qubo    Q = triu(U) + triu(U, 1)^T, U ~ integers[low, high] (inclusive)
  symmetric_normal_qubo same fold of a float32 standard-normal matrix
  maxcut_gnp            G(n, p) unit-weight MaxCut (max_cut.py:59-62 convention)
  tsp_int               integer TSP: cities ~ U[0,1)^2, distances * 1000 rounded,
                        penalty A = n_cities * max distance, structure of tsp.py:285-321
  batch_mix             4096-problem mix: even i MaxCut G(n, 0.5), n ~ U{16..32};
                        odd i integer TSP with 4 or 5 cities (16 / 25 variables)
  serving_requests      Poisson arrivals, n in 20..36 with P(n) ~ 2^-(n-20), half
                        symmetric int16 RandomQUBO in [-127, 127], half MaxCut G(n, 0.5)
All draws come from numpy's PCG64 with the given seed, in the order documented per
function, so every workload is reproducible.
"""
from __future__ import annotations

import numpy as np
from numpy.random import PCG64, Generator

from .just_Q import Just_Q
from .max_cut import MaxCut
from .tsp import distances_from_positions, qubo_from_distances


def symmetric_int_qubo(n: int, seed: int, low: int = -127, high: int = 127, dtype=np.int16) -> np.ndarray:
    rng = Generator(PCG64(seed))
    U = rng.integers(low, high + 1, (n, n), dtype=dtype)
    return np.triu(U) + np.triu(U, 1).T


def symmetric_normal_qubo(n: int, seed: int) -> np.ndarray:
    rng = Generator(PCG64(seed))
    G = rng.standard_normal((n, n), dtype=np.float32)
    return np.triu(G) + np.triu(G, 1).T


def maxcut_gnp(n: int, p: float = 0.5, rng=None, seed=None, dtype=np.int32) -> np.ndarray:
    return MaxCut.random_gnp(n, p, seed=seed, rng=rng, dtype=dtype).build_Q_matrix()


def tsp_int(n_cities: int, rng=None, seed=None) -> np.ndarray:
    if rng is None:
        rng = Generator(PCG64(seed))
    cities = rng.random((n_cities, 2))
    D = np.rint(1000.0 * distances_from_positions(cities[:, 0], cities[:, 1])).astype(np.int64)
    A = int(n_cities * D.max())
    return qubo_from_distances(D, penalty=A).astype(np.int32)


def batch_mix(count: int = 4096, seed: int = 0):
    """Returns (list of Just_Q problems, list of kinds)."""
    rng = Generator(PCG64(seed))
    probs, kinds = [], []
    for i in range(count):
        if i % 2 == 0:
            n = int(rng.integers(16, 33))
            probs.append(Just_Q(maxcut_gnp(n, 0.5, rng=rng)))
            kinds.append("maxcut")
        else:
            nc = int(rng.integers(4, 6))
            probs.append(Just_Q(tsp_int(nc, rng=rng)))
            kinds.append("tsp")
    return probs, kinds


def serving_n_distribution(n_min: int = 20, n_max: int = 36, ratio: float = 0.5) -> tuple:
    ns = np.arange(n_min, n_max + 1)
    w = ratio ** (ns - n_min)
    return ns, w / w.sum()


def serving_requests(count: int, rate_per_s: float, seed: int = 0, n_min: int = 20, n_max: int = 36):
    """Returns a list of (arrival_time_s, problem_Q ndarray, kind)."""
    rng = Generator(PCG64(seed))
    ns, pn = serving_n_distribution(n_min, n_max)
    t = 0.0
    out = []
    for _ in range(count):
        t += float(rng.exponential(1.0 / rate_per_s))
        n = int(rng.choice(ns, p=pn))
        if rng.random() < 0.5:
            U = rng.integers(-127, 128, (n, n), dtype=np.int16)
            Q = np.triu(U) + np.triu(U, 1).T
            kind = "random"
        else:
            Q = maxcut_gnp(n, 0.5, rng=rng, dtype=np.int16)
            kind = "maxcut"
        out.append((t, Q, kind))
    return out
