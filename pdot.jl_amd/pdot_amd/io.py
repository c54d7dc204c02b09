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


COST_KINDS = {"sqeuclid": 1, "l1": 2}


@dataclass
class PointCloudProblem:
    """A transport problem given by point clouds instead of a cost matrix
    (an extension beyond io.jl's OptimalTransportProblem): source points x
    (n x dim), target points y (m x dim), masses, and the cost
    C_ij = sum_d f(y_jd - x_id) with f(t) = t^2 ("sqeuclid") or |t| ("l1"),
    divided by its maximum if `normalize`.  The solver forms C in registers
    every pass; dense() builds the same matrix bit for bit."""
    source_points: np.ndarray
    target_points: np.ndarray
    source_distribution: np.ndarray
    target_distribution: np.ndarray
    cost: str = "sqeuclid"
    normalize: bool = False

    def __post_init__(self):
        self.source_points = np.ascontiguousarray(self.source_points, dtype=np.float64)
        self.target_points = np.ascontiguousarray(self.target_points, dtype=np.float64)
        self.source_distribution = np.ascontiguousarray(self.source_distribution, dtype=np.float64)
        self.target_distribution = np.ascontiguousarray(self.target_distribution, dtype=np.float64)
        if self.cost not in COST_KINDS:
            raise ValueError(f"unknown cost {self.cost!r}; expected one of {sorted(COST_KINDS)}")
        if self.source_points.ndim != 2 or self.target_points.ndim != 2 or \
                self.source_points.shape[1] != self.target_points.shape[1]:
            raise ValueError("points must be n x dim and m x dim arrays of the same dim")
        if not 1 <= self.source_points.shape[1] <= 3:
            raise ValueError("point dimension must be 1, 2 or 3")
        if self.source_distribution.shape != (self.num_row,) or self.target_distribution.shape != (self.num_col,):
            raise ValueError("distribution lengths do not match the point clouds")

    @property
    def num_row(self) -> int:
        return self.source_points.shape[0]

    @property
    def num_col(self) -> int:
        return self.target_points.shape[0]

    def dense(self) -> OptimalTransportProblem:
        x, y = self.source_points, self.target_points
        ct = np.zeros((y.shape[0], x.shape[0]), dtype=np.float64)
        for d in range(x.shape[1]):
            t = y[:, d][:, None] - x[:, d][None, :]
            ct += t * t if self.cost == "sqeuclid" else np.abs(t)
        C = ct.T
        if self.normalize:
            C = C / C.max()
        return OptimalTransportProblem(C, self.source_distribution, self.target_distribution)
