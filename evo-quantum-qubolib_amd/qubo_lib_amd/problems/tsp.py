"""Travelling salesman as a QUBO, mirroring qubo_lib's TSP (src/qubo_lib/problems/tsp.py:16-321).

Variable index city * n + time (tsp.py:226-235).  Tour cost couples consecutive times
(tsp.py:299-302); the penalty A = 2 * max(w) + 1 (tsp.py:305) multiplies the constraint
matrix C = -I + [same time, other city] + [same city, other time] (tsp.py:307-317); the
result is folded onto the upper triangle, triu(Q) + tril(Q, -1)^T (tsp.py:321), and cast
to the problem dtype (tsp.py:75).  Distances are computed in the cities' dtype exactly as
the reference computes them (tsp.py:273-282), so Q matches bit for bit.

The reference imports OR-Tools at module load for `calculate_ground` (tsp.py:11-13); here
OR-Tools is imported lazily, only by `calculate_ground`.
"""
from __future__ import annotations

import numpy as np
from numpy.random import PCG64, Generator

from ..abc import ProblemABC


def distances_from_positions(x_pos, y_pos) -> np.ndarray:
    """Pairwise Euclidean distances stored as float64 (tsp.py:273-282).  Each distance is
    evaluated on the coordinates' scalar type, element by element, as the reference
    evaluates it, so the float32 roundings are the same."""
    n = x_pos.shape[0]
    out = np.zeros((n, n))
    for i in range(n):
        for j in range(i + 1, n):
            out[i][j] = np.sqrt((x_pos[i] - x_pos[j]) ** 2 + (y_pos[i] - y_pos[j]) ** 2)
            out[j][i] = out[i][j]
    return out


def qubo_from_distances(w, penalty=None) -> np.ndarray:
    """Folded TSP QUBO for a distance matrix w (tsp.py:285-321).  `penalty` overrides the
    reference's A = 2 * max(w) + 1.  Integer w with an integer penalty stays integer."""
    w = np.asarray(w)
    n = w.shape[0]
    integer = np.issubdtype(w.dtype, np.integer) and (penalty is None or float(penalty).is_integer())
    acc = np.int64 if integer else np.float64
    N = n * n
    Q = np.zeros((N, N), dtype=acc)
    city = np.repeat(np.arange(n), n)          # index a = city * n + time
    time = np.tile(np.arange(n), n)
    # tour cost: Q[(i, p), (j, p + 1)] = w[i, j]
    for p in range(n - 1):
        rows = np.arange(n) * n + p
        cols = np.arange(n) * n + p + 1
        Q[np.ix_(rows, cols)] += w.astype(acc)
    A = (2 * w.max() + 1) if penalty is None else penalty
    A = acc(A) if integer else float(A)
    same_time = time[:, None] == time[None, :]
    same_city = city[:, None] == city[None, :]
    C = np.where(same_time & same_city, -1, 0) + np.where(same_time & ~same_city, 1, 0) \
        + np.where(same_city & ~same_time, 1, 0)
    Q = Q + A * C.astype(acc)
    return np.triu(Q) + np.tril(Q, k=-1).T


class TSP(ProblemABC):
    def __init__(self, N: int, dtype=np.float32):
        self.cities = None
        self.best_path = None
        super().__init__(N, dtype)

    @staticmethod
    def init_from_cities(cities, dtype=np.float32) -> "TSP":
        cities = np.asarray(cities)
        n_cities = cities.shape[0]
        obj = TSP(n_cities ** 2, dtype=dtype)
        obj.n_cities = n_cities
        obj.cities = cities.astype(dtype)
        return obj

    @staticmethod
    def init_random_instance(seed: int, n_cities: int, low=0, high=100, dtype=np.float32) -> "TSP":
        rng = Generator(PCG64(seed))
        cities = rng.uniform(low=low, high=high, size=(n_cities, 2)).astype(dtype)
        obj = TSP.init_from_cities(cities, dtype)
        obj.seed = seed
        obj.low, obj.high = low, high      # what a device-side build needs besides the seed
        return obj

    def build_Q_matrix(self, use_gpu: bool = False) -> np.ndarray:
        xs, ys = self.cities[:, 0], self.cities[:, 1]
        return self.qubo_from_positions(xs, ys).astype(self.dtype)

    @staticmethod
    def qubo_from_positions(x_pos, y_pos, lib=np) -> np.ndarray:
        return qubo_from_distances(distances_from_positions(x_pos, y_pos))

    @staticmethod
    def qubo_from_distances(w, lib=np) -> np.ndarray:
        return qubo_from_distances(w)

    # -- tour helpers (tsp.py:200-253) ------------------------------------------------
    @staticmethod
    def encode_path_to_x(path) -> np.ndarray:
        n = len(path)
        x = np.zeros(n * n)
        for t, c in enumerate(path):
            x[c * n + t] = 1
        return x

    @staticmethod
    def decode_x_to_path(wide_x) -> np.ndarray:
        wide_x = np.asarray(wide_x)
        N = wide_x.shape[0]
        n = int(round(N ** 0.5))
        assert n * n == N
        assert wide_x.sum() == n
        path = np.zeros(n, dtype=np.int32)
        for idx in np.nonzero(wide_x == 1)[0]:
            path[idx % n] = idx // n
        return path

    @staticmethod
    def get_path_length(path, xs, ys) -> float:
        length = 0
        for i in range(len(path) - 1):
            u, v = path[i], path[i + 1]
            length += ((xs[u] - xs[v]) ** 2 + (ys[u] - ys[v]) ** 2) ** 0.5
        return length

    def calculate_state_utility_metric(self, x) -> float:
        path = TSP.decode_x_to_path(x)
        return TSP.get_path_length(path, self.cities[:, 0], self.cities[:, 1])

    def calculate_utility_metric(self, states) -> np.ndarray:
        xs = states.get_wide_vectors()
        return np.array([self.calculate_state_utility_metric(x) for x in xs])

    def calculate_ground(self, use_gpu: bool = False):
        raise NotImplementedError("TSP.calculate_ground needs OR-Tools, which is not part of this build; "
                                  "for N <= ~48 solve the QUBO exactly with QBF instead")
