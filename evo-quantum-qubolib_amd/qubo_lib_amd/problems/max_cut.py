"""MaxCut as a QUBO, mirroring qubo_lib's MaxCut (src/qubo_lib/problems/max_cut.py:13-101).

For every edge (i, j, w): Q_ii -= w, Q_jj -= w, Q_ij += 2w (max_cut.py:59-62), so that
E(x) = -cut(x).  Named instances are read from a zip laid out like the reference's
well_known_instances/MaxCut.zip (instance/<name>/<name>.txt with 1-based "i j w" edge
lines, and best_solutions.csv with columns instance,vertices,edges,best).  Only the
small 'test' instance ships with this package; pass `zip_path` (or set the environment
variable QUBO_LIB_MAXCUT_ZIP) to use the full G-set archive.

Random G(n, p) graphs (not provided by the reference) are generated by `random_gnp`.
"""
from __future__ import annotations

import csv
import io
import os
import zipfile
from pathlib import Path

import numpy as np

from ..abc import ProblemABC

_BUNDLED = Path(__file__).resolve().parent / "well_known_instances"


def _zip_path(zip_path=None):
    if zip_path is not None:
        return Path(zip_path)
    env = os.environ.get("QUBO_LIB_MAXCUT_ZIP")
    return Path(env) if env else None


def _read_known(zip_path=None) -> dict:
    zp = _zip_path(zip_path)
    if zp is not None:
        with zipfile.ZipFile(zp) as z, z.open("best_solutions.csv") as f:
            return _parse_csv(io.TextIOWrapper(f, encoding="ascii"))
    with open(_BUNDLED / "best_solutions.csv", encoding="ascii") as f:
        return _parse_csv(f)


def _parse_csv(fh) -> dict:
    out = {}
    for row in csv.DictReader(fh):
        out[row["instance"].strip()] = (int(row["vertices"].strip()), int(row["best"].strip()))
    return out


def _edges_from_text(lines):
    it = iter(lines)
    num_vertices, num_edges = map(int, next(it).strip().split())
    edges = []
    for _ in range(num_edges):
        i, j, w = map(int, next(it).strip().split())
        edges.append((i - 1, j - 1, w))
    return num_vertices, edges


class MaxCut(ProblemABC):
    def __init__(self, instance_name: str, zip_path=None):
        known = _read_known(zip_path)
        if instance_name not in known:
            raise ValueError("Unknown instance name")
        self.instance_name = instance_name
        self.zip_path = None if zip_path is None else str(zip_path)
        self.edges = None
        N, best = known[instance_name]
        super().__init__(N)
        self.ground_E = -best
        self.ground_utility = self.ground_E

    @classmethod
    def from_edges(cls, N: int, edges, dtype=np.float32, name: str = "custom") -> "MaxCut":
        """Graph given as (i, j, w) triples with 0-based vertex indices."""
        obj = cls.__new__(cls)
        ProblemABC.__init__(obj, N, dtype)
        obj.instance_name = name
        obj.zip_path = None
        obj.edges = [(int(i), int(j), int(w)) for i, j, w in edges]
        return obj

    @classmethod
    def random_gnp(cls, n: int, p: float = 0.5, seed=None, rng=None, dtype=np.int32) -> "MaxCut":
        """Erdos-Renyi G(n, p) with unit weights: edge (u, v), u < v, iff rng.random() < p,
        drawn in row-major order of u then v."""
        if rng is None:
            rng = np.random.Generator(np.random.PCG64(seed))
        iu, iv = np.triu_indices(n, 1)
        keep = rng.random(iu.shape[0]) < p
        edges = [(int(u), int(v), 1) for u, v in zip(iu[keep], iv[keep])]
        obj = cls.from_edges(n, edges, dtype=dtype, name=f"gnp_{n}_{p}")
        obj.seed = seed
        return obj

    def _edges(self):
        if self.edges is not None:
            return self.N, self.edges
        zp = _zip_path(self.zip_path)
        member = f"instance/{self.instance_name}/{self.instance_name}.txt"
        if zp is not None:
            with zipfile.ZipFile(zp) as z, z.open(member) as f:
                return _edges_from_text(io.TextIOWrapper(f, encoding="ascii"))
        path = _BUNDLED / f"{self.instance_name}.txt"
        if not path.exists():
            raise FileNotFoundError(f"instance '{self.instance_name}' is not bundled; pass zip_path= "
                                    "or set QUBO_LIB_MAXCUT_ZIP to the MaxCut.zip archive")
        with open(path, encoding="ascii") as f:
            return _edges_from_text(f)

    def build_Q_matrix(self, use_gpu: bool = False) -> np.ndarray:
        num_vertices, edges = self._edges()
        Q = np.zeros((num_vertices, num_vertices), dtype=self.dtype)
        for i, j, w in edges:
            Q[i, i] -= w
            Q[j, j] -= w
            Q[i, j] += 2 * w
        assert Q.dtype == self.dtype
        return Q

    @staticmethod
    def init_random_instance(*args, **kwargs):
        raise NotImplementedError("not implemented yet; use MaxCut.random_gnp")

    @staticmethod
    def get_known_instances(zip_path=None) -> dict:
        return _read_known(zip_path)
