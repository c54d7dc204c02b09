"""Problem type (src/io.jl:1-5).  The device transfer (ot_cpu_to_gpu /
gpu_to_cpu!, io.jl:18-53) is DeviceSolver.upload/download in _lib.py."""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor
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
    every pass; dense() builds the same matrix bit for bit (the FP64-storage
    solve forms exactly it; FP32 storage multiplies by the reciprocal of the
    maximum instead of dividing, within about one FP64 ulp)."""
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
        C = point_cost_matrix(self.source_points, self.target_points, self.cost)
        if self.normalize:
            s = float(C.max())
            _parallel_columns(C.shape[1], lambda j0, j1: np.divide(C[:, j0:j1], s, out=C[:, j0:j1]))
        return OptimalTransportProblem(C, self.source_distribution, self.target_distribution)


def _parallel_columns(m: int, fn, min_cols: int = 64) -> None:
    """fn(j0, j1) over column ranges on host threads (numpy releases the GIL)."""
    workers = max(1, min(16, os.cpu_count() or 1, m // min_cols))
    if workers == 1:
        fn(0, m)
        return
    step = -(-m // (4 * workers))
    with ThreadPoolExecutor(workers) as ex:
        for f in [ex.submit(fn, j0, min(m, j0 + step)) for j0 in range(0, m, step)]:
            f.result()


def point_cost_matrix(x: np.ndarray, y: np.ndarray, cost: str) -> np.ndarray:
    """C_ij = sum_d f(y_jd - x_id), f(t) = t*t ("sqeuclid") or |t| ("l1"),
    summed from 0 in d order; column-major n x m, formed on host threads
    (each element by the same operations, so the result does not depend on the
    thread count)."""
    n, m = x.shape[0], y.shape[0]
    ct = np.zeros((m, n), dtype=np.float64)

    def rows(j0, j1):
        for d in range(x.shape[1]):
            t = y[j0:j1, d][:, None] - x[:, d][None, :]
            if cost == "sqeuclid":
                t *= t
            else:
                np.abs(t, out=t)
            ct[j0:j1] += t
    _parallel_columns(m, rows)
    return ct.T
