"""Synthetic optimal-transport instances for the five benchmark workloads.

The reference ships no generator, so instances come from a counter-based
SplitMix64 stream (seeded, platform independent): 53-bit uniforms,
Box-Muller normals, exp(N) lognormals and -log(U) Dirichlet(1) weights.  Cost
matrices are built in FP64, column-major (Fortran order), and masses are
normalised to sum 1 in FP64.

Workloads (n x m):
  uniform_sqeuclid   1000 x 1000: points U[0,1]^2, squared Euclidean cost,
                     masses U(0,1).
  gaussian_shift     10000 x 10000: source N(0, I2), target N((2,0), I2),
                     squared Euclidean cost divided by its maximum, masses U(0,1).
  l1_mixture_3d      30000 x 30000: source U[0,1]^3, target a 4-component
                     Gaussian mixture (sigma 0.1, means U[0,1]^3, equal
                     weights), L1 cost, lognormal(0,1) masses.
  batched_uniform    512 instances, n, m uniform integers in [64, 2048],
                     U[0,1]^2 clouds, squared Euclidean, U(0,1) masses.
  gmm_dirichlet      4096 x 4096: both sides 2-D Gaussian mixtures with 3-8
                     components (means U[0,1]^2, sigma U[0.02, 0.15],
                     Dirichlet(1) component weights), Dirichlet(1) masses,
                     squared Euclidean cost.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .io import OptimalTransportProblem, PointCloudProblem, point_cost_matrix

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


class Rng:
    """SplitMix64 stream: draw k consumes k counter values."""

    def __init__(self, seed: int):
        self.state = np.uint64(_mix(np.array([seed & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0])
        self.counter = 0

    def bits(self, k: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            idx = np.arange(self.counter + 1, self.counter + 1 + k, dtype=np.uint64)
            z = self.state + idx * _GAMMA
            out = _mix(z)
        self.counter += k
        return out

    def uniform(self, k: int) -> np.ndarray:
        """53-bit uniforms in [0, 1)."""
        return (self.bits(k) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def uniform_open(self, k: int) -> np.ndarray:
        """uniforms in (0, 1)."""
        return ((self.bits(k) >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)

    def normal(self, k: int) -> np.ndarray:
        """Box-Muller standard normals."""
        h = (k + 1) // 2
        u1 = self.uniform_open(h)
        u2 = self.uniform(h)
        r = np.sqrt(-2.0 * np.log(u1))
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:k]

    def integers(self, lo: int, hi: int, k: int) -> np.ndarray:
        """uniform integers in [lo, hi]."""
        span = hi - lo + 1
        return lo + np.minimum((self.uniform(k) * span).astype(np.int64), span - 1)

    def dirichlet1(self, k: int) -> np.ndarray:
        w = -np.log(self.uniform_open(k))
        return w / w.sum()


def normalized(w: np.ndarray) -> np.ndarray:
    return w / w.sum()


def sqeuclid_cost(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """C_ij = sum_d (x_id - y_jd)^2, returned column-major (n x m)."""
    return point_cost_matrix(x, y, "sqeuclid")


def l1_cost(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    return point_cost_matrix(x, y, "l1")


def uniform_sqeuclid_points(n: int = 1000, m: Optional[int] = None, seed: int = 1000) -> PointCloudProblem:
    m = n if m is None else m
    g = Rng(seed)
    x = g.uniform(2 * n).reshape(n, 2)
    y = g.uniform(2 * m).reshape(m, 2)
    a = normalized(g.uniform(n))
    b = normalized(g.uniform(m))
    return PointCloudProblem(x, y, a, b, "sqeuclid", normalize=False)


def uniform_sqeuclid(n: int = 1000, m: Optional[int] = None, seed: int = 1000) -> OptimalTransportProblem:
    return uniform_sqeuclid_points(n, m, seed).dense()


def gaussian_shift_points(n: int = 10000, m: Optional[int] = None, seed: int = 2000) -> PointCloudProblem:
    m = n if m is None else m
    g = Rng(seed)
    x = g.normal(2 * n).reshape(n, 2)
    y = g.normal(2 * m).reshape(m, 2)
    y[:, 0] += 2.0
    a = normalized(g.uniform(n))
    b = normalized(g.uniform(m))
    return PointCloudProblem(x, y, a, b, "sqeuclid", normalize=True)


def gaussian_shift(n: int = 10000, m: Optional[int] = None, seed: int = 2000) -> OptimalTransportProblem:
    return gaussian_shift_points(n, m, seed).dense()


def l1_mixture_3d_points(n: int = 30000, m: Optional[int] = None, seed: int = 3000) -> PointCloudProblem:
    m = n if m is None else m
    g = Rng(seed)
    x = g.uniform(3 * n).reshape(n, 3)
    means = g.uniform(12).reshape(4, 3)
    comp = g.integers(0, 3, m)
    y = means[comp] + 0.1 * g.normal(3 * m).reshape(m, 3)
    a = normalized(np.exp(g.normal(n)))
    b = normalized(np.exp(g.normal(m)))
    return PointCloudProblem(x, y, a, b, "l1", normalize=False)


def l1_mixture_3d(n: int = 30000, m: Optional[int] = None, seed: int = 3000) -> OptimalTransportProblem:
    return l1_mixture_3d_points(n, m, seed).dense()


def _gmm_points(g: Rng, k_pts: int) -> np.ndarray:
    K = int(g.integers(3, 8, 1)[0])
    means = g.uniform(2 * K).reshape(K, 2)
    sig = 0.02 + 0.13 * g.uniform(K)
    w = g.dirichlet1(K)
    comp = np.searchsorted(np.cumsum(w), g.uniform(k_pts), side="right")
    comp = np.minimum(comp, K - 1)
    return means[comp] + sig[comp][:, None] * g.normal(2 * k_pts).reshape(k_pts, 2)


@dataclass
class PointCloudInstance:
    """An instance with the point clouds it was built from (the learned warm
    start consumes coordinates and masses)."""
    x: np.ndarray
    y: np.ndarray
    problem: OptimalTransportProblem


def gmm_dirichlet_points(n: int = 4096, m: Optional[int] = None, seed: int = 5000):
    """The point clouds and masses (x, y, a, b) of gmm_dirichlet(n, m, seed),
    without forming the cost matrix."""
    m = n if m is None else m
    g = Rng(seed)
    x = _gmm_points(g, n)
    y = _gmm_points(g, m)
    a = g.dirichlet1(n)
    b = g.dirichlet1(m)
    return x, y, a, b


def gmm_dirichlet(n: int = 4096, m: Optional[int] = None, seed: int = 5000) -> PointCloudInstance:
    x, y, a, b = gmm_dirichlet_points(n, m, seed)
    return PointCloudInstance(x, y, OptimalTransportProblem(sqeuclid_cost(x, y), a, b))


def batched_uniform(count: int = 512, lo: int = 64, hi: int = 2048, seed: int = 4000) -> List[OptimalTransportProblem]:
    g = Rng(seed)
    ns = g.integers(lo, hi, count)
    ms = g.integers(lo, hi, count)
    out = []
    for k in range(count):
        n, m = int(ns[k]), int(ms[k])
        x = g.uniform(2 * n).reshape(n, 2)
        y = g.uniform(2 * m).reshape(m, 2)
        a = normalized(g.uniform(n))
        b = normalized(g.uniform(m))
        out.append(OptimalTransportProblem(sqeuclid_cost(x, y), a, b))
    return out


WORKLOADS = {
    "uniform_sqeuclid": uniform_sqeuclid,
    "gaussian_shift": gaussian_shift,
    "l1_mixture_3d": l1_mixture_3d,
    "gmm_dirichlet": gmm_dirichlet,
    "batched_uniform": batched_uniform,
}
