"""Problem type (src/io.jl:1-5).  The device transfer (ot_cpu_to_gpu /
gpu_to_cpu!, io.jl:18-53) is DeviceSolver.upload/download in _lib.py."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class OptimalTransportProblem:
    """io.jl:1-5: cost_matrix (n x m), source_distribution (n),
    target_distribution (m).  The matrix is kept column-major (Fortran
    order), the layout the C ABI takes without a copy."""
    cost_matrix: np.ndarray
    source_distribution: np.ndarray
    target_distribution: np.ndarray

    def __post_init__(self):
        self.cost_matrix = np.asfortranarray(self.cost_matrix, dtype=np.float64)
        self.source_distribution = np.ascontiguousarray(self.source_distribution, dtype=np.float64)
        self.target_distribution = np.ascontiguousarray(self.target_distribution, dtype=np.float64)
        if self.cost_matrix.ndim != 2:
            raise ValueError("cost_matrix must be a matrix")
        n, m = self.cost_matrix.shape
        if self.source_distribution.shape != (n,) or self.target_distribution.shape != (m,):
            raise ValueError("distribution lengths do not match the cost matrix")

    @property
    def num_row(self) -> int:
        return self.cost_matrix.shape[0]

    @property
    def num_col(self) -> int:
        return self.cost_matrix.shape[1]
