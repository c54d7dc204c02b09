"""What any learned starting iterate can gain, bounded on the CPU with the
oracle and exact optima (no GPU).

The learned warm start changes only the starting iterate of the same PDHG
solve (src/solver_gpu.jl:428-496 builds the zero start).  On instances of the
warm-start family (2-D Gaussian mixtures, Dirichlet masses, squared Euclidean
cost; here 128 x 128) the exact optimum (X*, u*, v*) of the transport LP,
from scipy's HiGHS, bounds every predictor:

* the exact optimum is a KKT point, so the solve stops at its first
  evaluation (0 iterations) at either primal weight;
* at the calibrated initial primal weight (0.13), optimal pairs perturbed by
  1% per entry, and the exact optimal duals with X = 0 (what a perfect dual
  predictor offers), take more iterations in total than the zero start.

So a predicted start helps only once it is already inside the tolerance: a
dual network cannot beat the zero start at the calibrated weight."""
from dataclasses import replace

import numpy as np
import pytest
from scipy import sparse
from scipy.optimize import linprog

import oracle as O
import pdot_amd as P
from pdot_amd import instances

SEEDS = (900200, 900201, 900202, 900207)
N = 128
CALIBRATED = 0.13


def exact_optimum(prob):
    """HiGHS on min <C, X> s.t. X 1 = a, X' 1 = b, X >= 0 (X column-major):
    the plan and the equality duals (u, v), with u_i + v_j <= C_ij."""
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    n, m = C.shape
    rows = sparse.kron(sparse.csr_matrix(np.ones((1, m))), sparse.eye(n))
    cols = sparse.kron(sparse.eye(m), sparse.csr_matrix(np.ones((1, n))))
    res = linprog(C.ravel(order="F"), A_eq=sparse.vstack([rows, cols]).tocsr(), b_eq=np.concatenate([a, b]),
                  bounds=(0, None), method="highs")
    assert res.status == 0
    y = res.eqlin.marginals
    return res.x.reshape((n, m), order="F"), y[:n], y[n:], res.fun


@pytest.fixture(scope="module")
def cases():
    out = []
    for seed in SEEDS:
        prob = instances.gmm_dirichlet(N, N, seed).problem
        X, u, v, f = exact_optimum(prob)
        # an exact optimum: feasible plan, feasible duals, zero gap (to rounding)
        assert X.min() >= -1e-12
        np.testing.assert_allclose(X.sum(axis=1), prob.source_distribution, atol=1e-12)
        assert (prob.cost_matrix - u[:, None] - v[None, :]).min() >= -1e-12
        assert abs(np.sum(prob.cost_matrix * X) - (u @ prob.source_distribution + v @ prob.target_distribution)) <= 1e-12
        out.append((seed, prob, X, u, v))
    return out


def _iterations(prob, weight, start=None, limit=2**31 - 1):
    params = replace(P.default_parameters(1e-4, iteration_limit=limit), primal_importance=weight)
    r = O.optimize(params, prob, warm_start=start, stats_cap=4)
    return r["iteration_count"], r["termination_reason"]


@pytest.mark.parametrize("weight", [1.0, CALIBRATED])
def test_exact_optimum_stops_at_the_first_evaluation(cases, weight):
    for seed, prob, X, u, v in cases:
        its, reason = _iterations(prob, weight, (u, v, X))
        assert (its, reason) == (0, int(P.TerminationReason.TERMINATION_REASON_OPTIMAL)), seed


def test_near_optimal_starts_lose_to_the_zero_start_at_the_calibrated_weight(cases):
    zero = perturbed = duals = 0
    for seed, prob, X, u, v in cases:
        z, reason = _iterations(prob, CALIBRATED)
        assert reason == int(P.TerminationReason.TERMINATION_REASON_OPTIMAL)
        zero += z
        cap = 16 * z  # a start this much slower has lost already
        g = np.random.default_rng(seed)
        d = 1e-2
        Xp = np.maximum(X * (1 + d * g.standard_normal(X.shape)), 0.0)
        up = u + d * np.abs(u).mean() * g.standard_normal(u.shape)
        vp = v + d * np.abs(v).mean() * g.standard_normal(v.shape)
        perturbed += _iterations(prob, CALIBRATED, (up, vp, Xp), limit=cap)[0]
        duals += _iterations(prob, CALIBRATED, (u, v, None), limit=cap)[0]
    print(f"iterations over {len(cases)} instances at weight {CALIBRATED}: zero start {zero}, "
          f"1%-perturbed optimum {perturbed}, exact duals with X = 0 {duals}")
    assert perturbed > zero
    assert duals > 2 * zero
