"""Bit-packed state vectors, mirroring qubo_lib's StateVectors (src/qubo_lib/state_vectors.py:5-80).

Bit j of a vector goes to element j // bits, bit j % bits, least significant first
(state_vectors.py:38-56), default element type uint8.
"""
from __future__ import annotations

import numpy as np


class StateVectors:
    SUPPORTED_TYPES = {np.uint8, np.uint16, np.uint32, np.uint64, np.int8, np.int16, np.int32, np.int64}

    def __init__(self, wide_vector_matrix: np.ndarray):
        wide = np.asarray(wide_vector_matrix)
        if wide.ndim != 2:
            raise ValueError("wide_vector_matrix must be 2-D (samples_cnt, N)")
        self.N = wide.shape[1]
        self.element_bits = 8
        self.bitwise_vectors = self.pack_wide_vectors_to_bitwise(wide)

    @classmethod
    def from_packed_vectors(cls, packed_vector_matrix: np.ndarray, N: int) -> "StateVectors":
        if packed_vector_matrix.dtype.type not in cls.SUPPORTED_TYPES:
            raise TypeError(f"Unsupported dtype: {packed_vector_matrix.dtype}. "
                            f"Supported types: {cls.SUPPORTED_TYPES}")
        obj = cls.__new__(cls)
        obj.bitwise_vectors = packed_vector_matrix
        obj.N = N
        obj.element_bits = packed_vector_matrix.dtype.itemsize * 8
        return obj

    def pack_wide_vectors_to_bitwise(self, wide_vectors: np.ndarray, target_dtype=np.uint8) -> np.ndarray:
        if target_dtype not in self.SUPPORTED_TYPES:
            raise TypeError(f"Unsupported target dtype: {target_dtype}. "
                            f"Supported types: {self.SUPPORTED_TYPES}")
        wide = np.asarray(wide_vectors) != 0
        count, n = wide.shape
        bits = np.dtype(target_dtype).itemsize * 8
        per = (n + bits - 1) // bits
        padded = np.zeros((count, per * bits), dtype=bool)
        padded[:, :n] = wide
        weights = (np.uint64(1) << np.arange(bits, dtype=np.uint64))
        grouped = padded.reshape(count, per, bits).astype(np.uint64)
        return (grouped * weights).sum(axis=2).astype(target_dtype)

    def get_wide_vectors(self) -> np.ndarray:
        if self.bitwise_vectors is None:
            return np.array([])
        packed = np.asarray(self.bitwise_vectors)
        count, per = packed.shape
        bits = self.element_bits
        j = np.arange(self.N)
        el = packed[:, j // bits].astype(np.uint64)
        return ((el >> (j % bits).astype(np.uint64)) & np.uint64(1)).astype(np.uint8)
