"""Learned warm start: the starting state of the same PDHG solve, learned
from solves of the problem family.

* Duals: a point-wise network with mean-pooled set context predicts the
  dual potentials (u, v) of a transport problem from its point clouds and
  masses; they set the starting iterate (X0 = 0, G0 = (C - u0) - v0).
* Primal weight: the solver state's initial primal weight (the reference's
  `primal_importance`, src/solver_gpu.jl:486-496, 1.0 by default).  The
  calibrated constant is the grid weight with the fewest total iterations over
  zero-start grid solves of training instances of the family
  (train_warm_start).  The network's weight head predicts a per-instance
  weight from the pooled context and the instance size, trained on each
  instance's best grid weight; it replaces the constant only if training
  found that it lowers the validation instances' iterations.

Zero-start PDHG with the caller's own parameters stays the fallback and the
solver's own KKT error decides: predicted duals whose first evaluation is not
better than the zero start's are not used, and a warm solve that does not
reach OPTIMAL is replaced by the plain zero-start solve.

The network runs on the GPU through pdot_mlp (libpdot_hip.so): bf16 MFMA
matrix products with FP32 accumulation, trained with Adam by the library's
own forward and backward kernels.

Features per point: (x1, x2, count * mass, side) with side +1 for sources
and -1 for targets.  Targets are the solver's duals normalised for the
(u + t, v - t) invariance of the dual objective: t = (a'u - b'v) / 2.
"""
from __future__ import annotations

import ctypes
import os
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib

_fp = ctypes.POINTER(ctypes.c_float)

MLP_SIGNATURES = {k: v for k, v in _lib.SIGNATURES.items() if k.startswith("pdot_mlp_")}

HIDDEN = 128
HEAD_WIDTH = 32
NUM_FEATURES = 4
# parameters before the weight head (a weights file without the head holds only these)
NUM_DUAL_PARAMS = 4 * 128 + 128 + 3 * 128 * 128 + 128 + 128 * 128 + 128 + 128 + 1
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warm_start_weights.npz")


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def features(x, y, a, b) -> np.ndarray:
    """N x 4 float32 rows: sources (x, n a_i, +1), then targets (y, m b_j, -1)."""
    x, y = np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)
    n, m = len(x), len(y)
    fs = np.column_stack([x[:, 0], x[:, 1], n * np.asarray(a), np.ones(n)])
    ft = np.column_stack([y[:, 0], y[:, 1], m * np.asarray(b), -np.ones(m)])
    return np.ascontiguousarray(np.vstack([fs, ft]), dtype=np.float32)


def normalize_duals(u, v, a, b):
    """(u - t, v + t) with t = (a'u - b'v)/2: same dual objective, balanced."""
    t = (np.dot(a, u) - np.dot(b, v)) / 2.0
    return u - t, v + t


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, in the pdot_mlp parameter order."""
    rng = np.random.default_rng(seed)

    def he(fan_in, shape):
        lim = np.sqrt(6.0 / fan_in)
        return rng.uniform(-lim, lim, size=shape)

    W = HIDDEN
    parts = [he(NUM_FEATURES, (NUM_FEATURES, W)), np.zeros(W), he(3 * W, (W, W)), he(3 * W, (W, W)),
             he(3 * W, (W, W)), np.zeros(W), he(W, (W, W)), np.zeros(W), he(W, (W,)) * 0.1, np.zeros(1)]
    return np.concatenate([p.ravel() for p in parts] + [init_head(seed)]).astype(np.float32)


def init_head(seed: int = 0) -> np.ndarray:
    """Weight-head parameters (Wh1a, Wh1b, wh1s, bh1, wh2, bh2): He-uniform
    first layer, small output layer, zero biases."""
    rng = np.random.default_rng(10_000 + seed)
    W, H = HIDDEN, HEAD_WIDTH
    lim = np.sqrt(6.0 / (2 * W + 1))
    parts = [rng.uniform(-lim, lim, (W, H)), rng.uniform(-lim, lim, (W, H)), rng.uniform(-lim, lim, H),
             np.zeros(H), rng.uniform(-0.1, 0.1, H), np.zeros(1)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


class WarmStartNet:
    """One pdot_mlp network on the GPU."""

    def __init__(self, params: Optional[np.ndarray] = None, device: int = 0, seed: int = 0):
        self.lib = _lib.load_library()
        for name, (res, args) in MLP_SIGNATURES.items():
            fn = getattr(self.lib, name)
            fn.restype, fn.argtypes = res, args
        h = ctypes.c_void_p()
        rc = self.lib.pdot_mlp_create(ctypes.byref(h), int(device))
        self.h = h
        if rc != _lib.PDOT_OK:
            msg = self.lib.pdot_mlp_last_error(h).decode() if h.value else str(rc)
            raise _lib.PdotError(f"pdot_mlp_create failed: {msg}")
        self.num_params = int(self.lib.pdot_mlp_num_params(h))
        self.set_params(init_params(seed) if params is None else params)

    def _check(self, rc, what):
        if rc != _lib.PDOT_OK:
            raise _lib.PdotError(f"{what} failed ({rc}): {self.lib.pdot_mlp_last_error(self.h).decode()}")

    def close(self):
        if self.h is not None and self.h.value:
            self.lib.pdot_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params) -> None:
        p, pp = _f32(params)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        self._check(self.lib.pdot_mlp_set_params(self.h, pp), "pdot_mlp_set_params")

    def get_params(self) -> np.ndarray:
        p = np.empty(self.num_params, dtype=np.float32)
        self._check(self.lib.pdot_mlp_get_params(self.h, p.ctypes.data_as(_fp)), "pdot_mlp_get_params")
        return p

    def forward(self, feats, ns: int, nt: int) -> np.ndarray:
        f, fp = _f32(feats)
        out = np.empty(ns + nt, dtype=np.float32)
        self._check(self.lib.pdot_mlp_forward(self.h, fp, int(ns), int(nt), out.ctypes.data_as(_fp)), "pdot_mlp_forward")
        return out

    def loss_grad(self, feats, ns: int, nt: int, target, want_grad: bool = False):
        f, fp = _f32(feats)
        t, tp = _f32(target)
        loss = ctypes.c_float()
        g = np.empty(self.num_params, dtype=np.float32) if want_grad else None
        self._check(self.lib.pdot_mlp_loss_grad(self.h, fp, int(ns), int(nt), tp, ctypes.byref(loss),
                                                g.ctypes.data_as(_fp) if g is not None else None), "pdot_mlp_loss_grad")
        return float(loss.value), g

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self._check(self.lib.pdot_mlp_adam_step(self.h, float(lr), float(beta1), float(beta2), float(eps)),
                    "pdot_mlp_adam_step")

    def gemm_probe(self, A, B, trans_a=False, trans_b=False) -> np.ndarray:
        A, ap = _f32(A)
        B, bp = _f32(B)
        M = A.shape[1] if trans_a else A.shape[0]
        K = A.shape[0] if trans_a else A.shape[1]
        N = B.shape[0] if trans_b else B.shape[1]
        C = np.empty((M, N), dtype=np.float32)
        self._check(self.lib.pdot_mlp_gemm_probe(self.h, ap, bp, C.ctypes.data_as(_fp), M, N, K, int(trans_a),
                                                 int(trans_b)), "pdot_mlp_gemm_probe")
        return C

    def predict_duals(self, x, y, a, b):
        """Warm-start duals (u0, v0) in float64."""
        n, m = len(x), len(y)
        out = self.forward(features(x, y, a, b), n, m).astype(np.float64)
        return out[:n], out[n:]

    def log_weight(self, feats, ns: int, nt: int) -> float:
        lw = ctypes.c_float()
        f, fp = _f32(feats)
        self._check(self.lib.pdot_mlp_predict_weight(self.h, fp, int(ns), int(nt), ctypes.byref(lw)),
                    "pdot_mlp_predict_weight")
        return float(lw.value)

    def predict_weight(self, x, y, a, b) -> float:
        """The per-instance initial primal weight of the weight head."""
        return float(np.exp(self.log_weight(features(x, y, a, b), len(x), len(y))))

    def weight_loss_grad(self, feats, ns: int, nt: int, target: float, train_trunk: bool = False,
                         want_grad: bool = False):
        """Squared error of the log weight; returns (loss, prediction, gradient or None)."""
        f, fp = _f32(feats)
        lw, loss = ctypes.c_float(), ctypes.c_float()
        g = np.empty(self.num_params, dtype=np.float32) if want_grad else None
        self._check(self.lib.pdot_mlp_weight_loss_grad(self.h, fp, int(ns), int(nt), float(target), int(train_trunk),
                                                       ctypes.byref(lw), ctypes.byref(loss),
                                                       g.ctypes.data_as(_fp) if g is not None else None),
                    "pdot_mlp_weight_loss_grad")
        return float(loss.value), float(lw.value), g


@dataclass
class Sample:
    feats: np.ndarray
    ns: int
    nt: int
    target: np.ndarray


def make_sample(x, y, a, b, u, v) -> Sample:
    un, vn = normalize_duals(u, v, a, b)
    return Sample(features(x, y, a, b), len(x), len(y), np.concatenate([un, vn]).astype(np.float32))


def train(net: WarmStartNet, samples: Sequence[Sample], epochs: int = 10, lr: float = 1e-3, seed: int = 0,
          lr_decay: float = 0.97, log=None) -> List[float]:
    """Adam over the samples (one instance per step), shuffled per epoch;
    returns the mean training loss per epoch."""
    rng = np.random.default_rng(seed)
    history = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = 0.0
        for i in order:
            s = samples[i]
            loss, _ = net.loss_grad(s.feats, s.ns, s.nt, s.target)
            net.adam_step(lr)
            tot += loss
        history.append(tot / len(samples))
        if log:
            log(f"epoch {ep}: loss {history[-1]:.4e} lr {lr:.2e}")
        lr *= lr_decay
    return history


def evaluate_loss(net: WarmStartNet, samples: Sequence[Sample]) -> float:
    return float(np.mean([net.loss_grad(s.feats, s.ns, s.nt, s.target)[0] for s in samples]))


def save_params(path: str, params: np.ndarray, meta: dict) -> None:
    np.savez(path, params=np.asarray(params, dtype=np.float32),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_params(path: str = DEFAULT_WEIGHTS) -> np.ndarray:
    """The network parameters; a file written before the weight head existed
    gets a freshly initialised head (has_weight_head() is then False)."""
    with np.load(path, allow_pickle=False) as z:
        p = z["params"].astype(np.float32)
    if p.size == NUM_DUAL_PARAMS:
        p = np.concatenate([p, init_head()])
    return p


def load_meta(path: str = DEFAULT_WEIGHTS) -> dict:
    """Everything stored beside the network parameters (training summary and
    the calibrated initial primal weight)."""
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files if k != "params"}


def load_initial_primal_weight(path: str = DEFAULT_WEIGHTS) -> Optional[float]:
    """The calibrated initial primal weight, or None if the weights file has none."""
    meta = load_meta(path)
    return float(meta["initial_primal_weight"]) if "initial_primal_weight" in meta else None


def has_weight_head(path: str = DEFAULT_WEIGHTS) -> bool:
    """Whether the weights file carries a trained per-instance weight head."""
    return "weight_head" in load_meta(path)


def weight_head_selected(path: str = DEFAULT_WEIGHTS) -> bool:
    """Whether training selected the head over the calibrated constant (it
    lowered the validation instances' iterations)."""
    meta = load_meta(path)
    return bool(int(meta["weight_head_selected"])) if "weight_head_selected" in meta else False


def zero_start_relative_kkt(problem) -> float:
    """Relative KKT error of the zero start at its first evaluation, in closed
    form: X = 0, u = v = 0 (so G = 0) leave only the primal residual
    ||[a; b]||_2, relative to 1 + ||[a; b]||_2 (src/utils_gpu.jl:91-124)."""
    q = float(np.sqrt(np.dot(problem.source_distribution, problem.source_distribution)
                      + np.dot(problem.target_distribution, problem.target_distribution)))
    return q / (1.0 + q)


@dataclass
class WarmStartInfo:
    """How solve_warm started the returned solve."""
    used_duals: bool                 # the predicted duals passed the KKT acceptance test
    primal_weight: Optional[float]   # the initial primal weight of the returned solve (None: the caller's)
    fell_back: bool                  # the warm solve missed OPTIMAL and the plain solve was returned
    predict_sec: float
    probe_sec: float = 0.0           # the KKT evaluation of the predicted duals


def learned_weight(net: WarmStartNet, x, y, a, b, path: str = DEFAULT_WEIGHTS) -> Optional[float]:
    """The learned initial primal weight: the weight head's prediction when
    training selected the head, else the calibrated constant (None if neither
    exists)."""
    if weight_head_selected(path):
        return net.predict_weight(x, y, a, b)
    return load_initial_primal_weight(path)


def first_evaluation_kkt(params, dev: _lib.DeviceSolver, u0, v0) -> float:
    """Relative KKT error of the starting state (X = 0, duals (u0, v0)) at a
    solve's first evaluation (src/solver_gpu.jl:529-537, utils_gpu.jl:68-129),
    from one device evaluation: the same quantity optimize() reports in its
    first IterationStats, without running or downloading a solve."""
    from . import solver as S
    from .solve_log import PointType
    ot_cache = dev.cached_ot_info()
    policy, re, ge, step_size, weight = S.initial_step_and_weight(params, dev.n, dev.m, ot_cache)
    dev.set_start(u0, v0, policy, re, ge, step_size, weight)
    cur, _, _ = dev.eval()
    crit = params.termination_criteria
    ci = S.convergence_information(cur, ot_cache, crit.eps_optimal_absolute / crit.eps_optimal_relative,
                                   PointType.POINT_TYPE_AVERAGE_ITERATE)
    return S.relative_kkt_error(ci)


def solve_warm(params, problem, x, y, net: WarmStartNet, *, storage=_lib.STORAGE_F64, fallback: bool = True,
               primal_weight: Optional[float] = None, use_calibrated_weight: bool = True,
               solver: Optional[_lib.DeviceSolver] = None, profile=None):
    """optimize() from the learned starting state, with the plain zero-start
    solve as the fallback and the solver's own KKT check deciding:

    1. the predicted duals are evaluated (the relative KKT error of a solve's
       first evaluation); if it is not below the zero start's, the solve
       starts from zero duals;
    2. the solve starts with the learned initial primal weight (the weight
       head's per-instance prediction if training selected it, else the
       calibrated constant; `primal_weight` overrides it and
       use_calibrated_weight=False keeps the caller's);
    3. a warm solve that does not reach OPTIMAL is replaced by the plain
       zero-start solve with the caller's parameters.

    solver: a DeviceSolver with the problem already uploaded (reused, not
    closed); otherwise one is created for the call.  profile: a SolveProfile
    for the returned solve.  Returns (output, info: WarmStartInfo)."""
    from dataclasses import replace
    from . import solver as S
    from .solve_log import TerminationReason
    t0 = time.time()
    u0, v0 = net.predict_duals(x, y, problem.source_distribution, problem.target_distribution)
    if primal_weight is None and use_calibrated_weight:
        primal_weight = learned_weight(net, x, y, problem.source_distribution, problem.target_distribution)
    pred = time.time() - t0
    warm_params = params if primal_weight is None else replace(params, primal_importance=float(primal_weight))
    n, m = problem.cost_matrix.shape
    dev = solver
    if dev is None:
        dev = _lib.DeviceSolver(n, m, storage=storage)
    elif (dev.n, dev.m) != (n, m):
        raise ValueError("solver was created for another problem size")
    try:
        if solver is None:
            dev.upload(problem.cost_matrix, problem.source_distribution, problem.target_distribution)
        t0 = time.time()
        warm_kkt = first_evaluation_kkt(params, dev, u0, v0)
        probe = time.time() - t0
        used = bool(warm_kkt < zero_start_relative_kkt(problem))
        out = S.optimize(warm_params, problem, storage=storage, warm_start=(u0, v0) if used else None, solver=dev,
                         profile=profile)
        if out.termination_reason == TerminationReason.TERMINATION_REASON_OPTIMAL or not fallback:
            return out, WarmStartInfo(used, primal_weight, False, pred, probe)
        return (S.optimize(params, problem, storage=storage, solver=dev, profile=profile),
                WarmStartInfo(False, None, True, pred, probe))
    finally:
        if solver is None:
            dev.close()
