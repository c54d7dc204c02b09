"""Pins the CPU oracle (oracle/pdot_oracle.c) before it is trusted as the
checker of the HIP path.

The reference holds no tests or golden vectors, so the pins are:
  * known answers derived by hand from src/solver_gpu.jl: from the zero
    start, pass 1 gives Delta = 0, u = a, v = b and interaction 0, so the step
    size goes 1 -> 1.5; iteration 2 restarts to the average (artificially,
    with omega unchanged) and pass 2 gives X = 1.5 max(a_i + b_j - C_ij, 0);
  * LP optima from independent solvers (scipy HiGHS, linear_sum_assignment),
    committed in tests/golden/lp_optima.json;
  * closed forms: separable costs (every feasible plan is optimal), n = 1
    (the plan is forced), 1-D convex costs (the monotone coupling);
  * invariants: the running average equals sum_k w_k X_k / sum_k w_k;
  * the termination/limit semantics of termination.jl and solver_gpu.jl.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
import pdot_amd as P
from pdot_amd import instances
from pdot_amd.solve_log import TerminationReason as TR

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "lp_optima.json")))


def fixture_problem(case):
    out = instances.WORKLOADS[case["generator"]](case["n"], case["m"], seed=case["seed"])
    prob = out.problem if hasattr(out, "problem") else out
    assert prob.cost_matrix.sum() == pytest.approx(case["cost_checksum"], rel=1e-12)
    return prob


def find_case(gen, n, m, seed):
    for c in FIX["lp"]:
        if (c["generator"], c["n"], c["m"], c["seed"]) == (gen, n, m, seed):
            return c
    raise KeyError((gen, n, m, seed))


@pytest.mark.parametrize("mode", [O.LITERAL, O.FUSED])
def test_known_answer_first_two_iterations(mode):
    prob = instances.uniform_sqeuclid(37, 23, seed=5)
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    params = P.default_parameters(1e-4, iteration_limit=3)
    r = O.optimize(params, prob, mode=mode, trace_cap=10)
    # pass 1: interaction 0 -> limit inf -> eta = (1 + 1/2) * 1
    assert r["eta_trace"][0] == 1.5
    st = r["stats"]
    assert [s["iteration_number"] for s in st] == [1, 2, 3]
    # iteration 1 evaluates the zero start: P = ||[a;b]||, D = 0, objectives 0
    assert st[0]["primal_objective"] == 0.0 and st[0]["dual_objective"] == 0.0
    assert st[0]["l2_dual_residual"] == 0.0
    assert st[0]["l2_primal_residual"] == pytest.approx(math.sqrt(a @ a + b @ b), rel=1e-15)
    # iteration 2 evaluates (X=0, u=a, v=b): dual objective a'a + b'b
    assert st[1]["dual_objective"] == pytest.approx(a @ a + b @ b, rel=1e-15)
    assert st[1]["restart_used"] == int(P.RestartChoice.RESTART_CHOICE_RESTART_TO_AVERAGE)
    assert st[1]["primal_weight"] == 1.0 and st[2]["primal_weight"] == 1.0
    # after pass 2 (eta 1.5, omega 1) the plan is 1.5 max(a_i + b_j - C_ij, 0); the
    # iteration-3 evaluation reports ||X||_F of the average = that plan
    X2 = np.maximum(-(1.5 * ((C - a[:, None]) - b[None, :])), 0.0)
    assert st[2]["l2_primal_variable"] == pytest.approx(np.sqrt(np.sum(X2 * X2)), rel=1e-14)
    assert st[2]["primal_objective"] == pytest.approx(np.sum(C * X2), rel=1e-14)


@pytest.mark.parametrize("key", [("uniform_sqeuclid", 100, 80, 7), ("uniform_sqeuclid", 1, 50, 3),
                                 ("uniform_sqeuclid", 40, 1, 4), ("gaussian_shift", 400, 400, 2000)])
def test_oracle_reaches_lp_optimum(key):
    case = find_case(*key)
    prob = fixture_problem(case)
    r = O.optimize(P.default_parameters(1e-8, iteration_limit=300000), prob, mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_OPTIMAL)
    pobj = r["stats"][-1]["primal_objective"]
    assert abs(pobj - case["optimum"]) <= 5e-8 * (1 + abs(case["optimum"]))


def test_oracle_matches_assignment_optimum():
    case = FIX["assignment"][0]
    rng = np.random.default_rng(case["seed"])
    n = case["n"]
    x, y = rng.random((n, 2)), rng.random((n, 2))
    C = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    assert C.sum() == pytest.approx(case["cost_checksum"], rel=1e-12)
    prob = P.OptimalTransportProblem(C, np.full(n, 1 / n), np.full(n, 1 / n))
    r = O.optimize(P.default_parameters(1e-8, iteration_limit=300000), prob, mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_OPTIMAL)
    assert abs(r["stats"][-1]["primal_objective"] - case["optimum"]) <= 5e-8


def test_separable_cost_closed_form():
    rng = np.random.default_rng(3)
    n, m = 30, 45
    f, g = rng.random(n), rng.random(m)
    C = f[:, None] + g[None, :]
    a = rng.random(n); a /= a.sum()
    b = rng.random(m); b /= b.sum()
    r = O.optimize(P.default_parameters(1e-8), P.OptimalTransportProblem(C, a, b), mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_OPTIMAL)
    assert r["stats"][-1]["primal_objective"] == pytest.approx(a @ f + b @ g, abs=1e-7)


def test_single_source_plan_is_forced():
    rng = np.random.default_rng(4)
    m = 60
    C = rng.random((1, m))
    b = rng.random(m); b /= b.sum()
    r = O.optimize(P.default_parameters(1e-8), P.OptimalTransportProblem(C, np.ones(1), b), mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_OPTIMAL)
    np.testing.assert_allclose(r["X"][0], b, atol=1e-7)


def northwest_corner_value(x, y, a, b):
    """Monotone coupling of sorted 1-D points: optimal for convex costs."""
    ia, ib = np.argsort(x), np.argsort(y)
    aa, bb = a[ia].copy(), b[ib].copy()
    i = j = 0
    val = 0.0
    while i < len(aa) and j < len(bb):
        t = min(aa[i], bb[j])
        val += t * (x[ia[i]] - y[ib[j]]) ** 2
        aa[i] -= t
        bb[j] -= t
        if aa[i] <= 1e-300:
            i += 1
        if j < len(bb) and bb[j] <= 1e-300:
            j += 1
    return val


def test_one_dimensional_convex_cost_is_monotone_coupling():
    rng = np.random.default_rng(8)
    n, m = 40, 33
    x, y = rng.random(n), rng.random(m) + 0.3
    a = rng.random(n); a /= a.sum()
    b = rng.random(m); b /= b.sum()
    C = (x[:, None] - y[None, :]) ** 2
    r = O.optimize(P.default_parameters(1e-8, iteration_limit=300000), P.OptimalTransportProblem(C, a, b),
                   mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_OPTIMAL)
    assert r["stats"][-1]["primal_objective"] == pytest.approx(northwest_corner_value(x, y, a, b), abs=1e-7)


def test_literal_and_fused_modes_agree():
    prob = instances.uniform_sqeuclid(120, 90, seed=21)
    params = P.default_parameters(1e-4)
    lit = O.optimize(params, prob, mode=O.LITERAL)
    fus = O.optimize(params, prob, mode=O.FUSED)
    assert lit["termination_reason"] == fus["termination_reason"] == int(TR.TERMINATION_REASON_OPTIMAL)
    # the first ten evaluations agree to rounding
    for sl, sf in zip(lit["stats"][:10], fus["stats"][:10]):
        assert sl["restart_used"] == sf["restart_used"]
        for k in ("primal_objective", "dual_objective", "l2_primal_residual", "l2_dual_residual", "step_size",
                  "primal_weight"):
            assert sl[k] == pytest.approx(sf[k], rel=1e-11, abs=1e-15), k
    pl, pf = lit["stats"][-1]["primal_objective"], fus["stats"][-1]["primal_objective"]
    assert abs(pl - pf) <= 10 * 1e-4 * (1 + abs(pl))


def test_weighted_average_invariant():
    """X_avg after k passes equals sum_k w_k X_k / sum_k w_k (weights = eta)."""
    prob = instances.uniform_sqeuclid(20, 17, seed=2)
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    n, m = C.shape
    X, Xa = np.zeros((n, m), order="F"), np.zeros((n, m), order="F")
    u, v, r, c = np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(m)
    ua, va, ra, ca = np.zeros(n), np.zeros(m), np.zeros(n), np.zeros(m)
    S, acc, etas = 0.0, np.zeros((n, m)), [1.0, 1.5, 0.7, 2.0, 0.3, 1.1]
    gz = 1
    for eta in etas:
        O.fused_pass(C, a, b, X, u, v, r, c, gz, eta, 1.3, Xa, ua, va, ra, ca, S)
        gz = 0
        acc += eta * X
        S += eta
        np.testing.assert_allclose(Xa, acc / S, rtol=1e-13, atol=1e-16)
        np.testing.assert_allclose(r, X.sum(axis=1), rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(c, X.sum(axis=0), rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("limit,count", [(5, 4), (10, 9), (11, 11), (200, 200)])
def test_iteration_limit_semantics(limit, count):
    """Evaluations run at iterations 1..10, every 128 and at limit+1
    (solver_gpu.jl:529-532); termination at iteration_number >= limit returns
    iteration-1 (termination.jl:156, solver_gpu.jl:661)."""
    prob = instances.uniform_sqeuclid(50, 40, seed=1)
    r = O.optimize(P.default_parameters(1e-12, iteration_limit=limit), prob, mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_ITERATION_LIMIT)
    assert r["iteration_count"] == count


def test_kkt_pass_and_time_limits():
    prob = instances.uniform_sqeuclid(50, 40, seed=1)
    p = P.default_parameters(1e-12)
    p.termination_criteria.kkt_matrix_pass_limit = 300.0
    r = O.optimize(p, prob, mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT)
    assert r["stats"][-1]["cumulative_kkt_matrix_passes"] >= 300.0
    p = P.default_parameters(1e-12, time_sec_limit=1e-9)
    r = O.optimize(p, prob, mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_TIME_LIMIT)
    assert r["iteration_count"] == 0


def test_nan_cost_runs_to_the_iteration_limit():
    """A NaN cost makes the interaction NaN, which the acceptance test reads as
    'not > 0' (infinite step limit, solver_gpu.jl:338-347): the reference keeps
    iterating on NaN iterates until a limit, and so does the oracle."""
    prob = instances.uniform_sqeuclid(10, 10, seed=1)
    C = prob.cost_matrix.copy()
    C[3, 4] = np.nan
    r = O.optimize(P.default_parameters(1e-4, iteration_limit=40),
                   P.OptimalTransportProblem(C, prob.source_distribution, prob.target_distribution), mode=O.FUSED)
    assert r["termination_reason"] == int(TR.TERMINATION_REASON_ITERATION_LIMIT)
    assert math.isnan(r["stats"][-1]["primal_objective"])


def test_invalid_parameters_rejected():
    prob = instances.uniform_sqeuclid(10, 10, seed=1)
    p = P.default_parameters(1e-4)
    p.primal_importance = -1.0
    with pytest.raises(RuntimeError):
        O.optimize(p, prob)
    with pytest.raises(AssertionError):
        P.construct_restart_parameters(P.ADAPTIVE_KKT, P.KKT_GREEDY, 1000, 0.36, 0.9, 0.1, 0.5)


def test_cached_ot_info_matches_numpy():
    prob = instances.uniform_sqeuclid(33, 44, seed=9)
    o = O.cached_ot_info(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    h = P.cached_ot_info(prob)
    np.testing.assert_allclose(o, [h.l_inf_norm_primal_linear_objective, h.l_inf_norm_primal_right_hand_side,
                                   h.l2_norm_primal_linear_objective, h.l2_norm_primal_right_hand_side], rtol=1e-14)


def test_rounding_is_feasible_and_close():
    """rounding (utils_gpu.jl:296-318) maps a near-feasible plan onto the
    transport polytope exactly (up to rounding)."""
    prob = instances.uniform_sqeuclid(30, 20, seed=3)
    r = O.optimize(P.default_parameters(1e-4), prob, mode=O.FUSED)
    X = O.rounding(prob.source_distribution, prob.target_distribution, r["X"])
    np.testing.assert_allclose(X.sum(axis=1), prob.source_distribution, atol=1e-15)
    np.testing.assert_allclose(X.sum(axis=0), prob.target_distribution, atol=1e-15)
    assert X.min() >= -1e-15  # the rank-1 correction can leave rounding-level negatives


def test_rounding_known_answer_by_hand():
    """2 x 2 case worked by hand through src/utils_gpu.jl:296-318:
    X = [.4 .2; 0 .1], a = b = (.5, .5): s = (5/6, 1), t = (1, 1),
    e_r = (0, .4)/.4, e_c = (1/6, 7/30), Y = [1/3 1/6; 1/6 1/3]."""
    X = np.array([[0.4, 0.2], [0.0, 0.1]], order="F")
    Y = O.rounding(np.array([0.5, 0.5]), np.array([0.5, 0.5]), X)
    np.testing.assert_allclose(Y, [[1 / 3, 1 / 6], [1 / 6, 1 / 3]], rtol=0, atol=2e-16)


def test_rounding_leaves_a_feasible_plan_unchanged():
    """A feasible plan with dyadic entries (exact sums): s = t = 1 and the
    residuals are exactly 0, so the output is bit-identical."""
    rng = np.random.default_rng(4)
    a = rng.integers(1, 64, 37).astype(float)
    b = rng.integers(1, 64, 23).astype(float)
    a /= 2.0 ** np.ceil(np.log2(a.sum()))
    a[0] += 1.0 - a.sum()
    b /= 2.0 ** np.ceil(np.log2(b.sum()))
    b[0] += 1.0 - b.sum()
    X = np.asfortranarray(np.outer(a, b))
    assert np.array_equal(X.sum(axis=1), a) and np.array_equal(X.sum(axis=0), b)
    assert np.array_equal(O.rounding(a, b, X), X)


def test_rounding_error_bound_and_zero_rows():
    """Altschuler, Weed, Rigollet (2017), Lemma 7: the rounded plan Y is
    feasible and ||Y - X||_1 <= 2 (||r(X) - a||_1 + ||c(X) - b||_1).  A row
    and a column of zeros (0/0 -> 0 scaling) and masses above the plan's sums
    (ratio capped at 1) are included."""
    rng = np.random.default_rng(5)
    for n, m in ((1, 9), (9, 1), (17, 31), (64, 50)):
        a = rng.random(n); a /= a.sum()
        b = rng.random(m); b /= b.sum()
        X = np.asfortranarray(rng.random((n, m)) * rng.choice([0.5, 2.0], size=(n, 1)) / (n * m))
        X[n // 2, :] = 0.0
        X[:, m // 2] = 0.0
        Y = O.rounding(a, b, X)
        np.testing.assert_allclose(Y.sum(axis=1), a, atol=1e-15)
        np.testing.assert_allclose(Y.sum(axis=0), b, atol=1e-15)
        bound = 2 * (np.abs(X.sum(axis=1) - a).sum() + np.abs(X.sum(axis=0) - b).sum())
        assert np.abs(Y - X).sum() <= bound * (1 + 1e-12)
        assert Y.min() >= -1e-15
