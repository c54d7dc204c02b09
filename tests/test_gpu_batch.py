"""GPU tests of the batched form (pdot_batch): many problems of uneven size
advanced together must give, instance by instance, what the single-problem
semantics give.  Per-instance parity is pinned like the single path's: the
first ten evaluations against the oracle to 1e-9, whole solves by results
(termination, recomputed KKT, objective against the HiGHS optima)."""
import json
import os

import numpy as np
import pytest

import oracle as O
import pdot_amd as P
from pdot_amd import _lib, instances
from pdot_amd.batch import autotune, optimize_batch

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "lp_optima.json")))

EVAL_FIELDS = ("primal_objective", "dual_objective", "l2_primal_residual", "l_inf_primal_residual",
               "l2_dual_residual", "l_inf_dual_residual", "l2_primal_variable", "relative_optimality_gap")


def mixed_problems():
    sizes = [(64, 64), (300, 257), (1, 40), (50, 1), (1100, 70), (130, 1500), (700, 700)]
    return [instances.uniform_sqeuclid(n, m, seed=10 * k + 1) for k, (n, m) in enumerate(sizes)]


def recomputed_kkt(prob, X, u, v):
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    r, c = X.sum(axis=1), X.sum(axis=0)
    P2 = np.sqrt(np.sum((a - r) ** 2) + np.sum((b - c) ** 2))
    D2 = np.sqrt(np.sum(np.minimum(C - u[:, None] - v[None, :], 0.0) ** 2))
    pobj, dobj = float(np.sum(C * X)), float(a @ u + b @ v)
    return max(P2 / (1 + np.sqrt(a @ a + b @ b)), D2 / (1 + np.sqrt(np.sum(C * C))),
               abs(pobj - dobj) / (1 + abs(pobj) + abs(dobj))), pobj


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_batch_first_ten_evaluations_match_oracle(storage):
    probs = mixed_problems()
    params = P.default_parameters(1e-4, iteration_limit=10)
    outs = optimize_batch(params, probs, storage=storage, use_graphs=False)
    for k, (prob, out) in enumerate(zip(probs, outs)):
        ref = O.optimize(params, prob, mode=O.FUSED)
        assert int(out.termination_reason) == ref["termination_reason"], k
        assert out.iteration_count == ref["iteration_count"], k
        assert len(out.iteration_stats) == ref["n_stats"], k
        if storage == _lib.STORAGE_F32:
            continue  # FP32 storage changes the arithmetic; pinned by results below
        for s_dev, s_ref in zip(out.iteration_stats, ref["stats"]):
            assert int(s_dev.restart_used) == s_ref["restart_used"]
            np.testing.assert_allclose(s_dev.step_size, s_ref["step_size"], rtol=1e-9)
            np.testing.assert_allclose(s_dev.primal_weight, s_ref["primal_weight"], rtol=1e-9)
            ci = s_dev.convergence_information[0]
            for f in EVAL_FIELDS:
                np.testing.assert_allclose(getattr(ci, f), s_ref[f], rtol=1e-9, atol=1e-15, err_msg=f"{k} {f}")
        np.testing.assert_allclose(out.primal_solution, ref["X"], rtol=1e-9, atol=1e-15)


def fixture_batch():
    cases = [c for c in FIX["lp"] if c["n"] * c["m"] <= 1_000_000]
    probs = []
    for c in cases:
        out = instances.WORKLOADS[c["generator"]](c["n"], c["m"], seed=c["seed"])
        probs.append(out.problem if hasattr(out, "problem") else out)
    return cases, probs


@pytest.mark.parametrize("storage,tol", [(_lib.STORAGE_F32, 1e-4), (_lib.STORAGE_F64, 1e-6)])
def test_batch_results_match_lp_optima(storage, tol):
    cases, probs = fixture_batch()
    outs = optimize_batch(P.default_parameters(tol, iteration_limit=400000), probs, storage=storage)
    for c, prob, out in zip(cases, probs, outs):
        assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL, c
        ci = out.iteration_stats[-1].convergence_information[0]
        assert P.relative_kkt_error(ci) < tol
        kkt, pobj = recomputed_kkt(prob, out.primal_solution, out.dual_source_solution, out.dual_target_solution)
        slack = 1e-6 * np.abs(prob.cost_matrix).max() if storage == _lib.STORAGE_F32 else 0.0
        assert kkt < 1.01 * tol + slack, c
        assert abs(pobj - c["optimum"]) <= 10 * tol * (1 + abs(c["optimum"])), c


def test_batch_graph_replays_match_eager_bitwise():
    """The evaluation interval captured as a hipGraph and replayed many times
    (with evaluations, restarts and host synchronisations between replays)
    gives bit-identical results to plain stream launches."""
    probs = mixed_problems()[:5]
    params = P.default_parameters(1e-4, iteration_limit=400000)
    eager = optimize_batch(params, probs, storage=_lib.STORAGE_F32, use_graphs=False)
    graph = optimize_batch(params, probs, storage=_lib.STORAGE_F32, use_graphs=True)
    for e, g in zip(eager, graph):
        assert e.iteration_count == g.iteration_count
        assert e.termination_reason == g.termination_reason
        assert np.array_equal(e.primal_solution, g.primal_solution)
        assert np.array_equal(e.dual_source_solution, g.dual_source_solution)
        for se, sg in zip(e.iteration_stats, g.iteration_stats):
            assert se.convergence_information[0].primal_objective == sg.convergence_information[0].primal_objective


def test_single_graph_replays_match_eager_bitwise():
    prob = instances.uniform_sqeuclid(400, 333, seed=4)
    params = P.default_parameters(1e-4, iteration_limit=400000)
    e = P.optimize(params, prob, use_graphs=False)
    g = P.optimize(params, prob, use_graphs=True)
    assert e.iteration_count == g.iteration_count
    assert np.array_equal(e.primal_solution, g.primal_solution)


def test_batch_matches_single_instance_results():
    probs = mixed_problems()
    params = P.default_parameters(1e-4, iteration_limit=400000)
    outs = optimize_batch(params, probs, storage=_lib.STORAGE_F64)
    for prob, out in zip(probs, outs):
        single = P.optimize(params, prob)
        assert out.termination_reason == single.termination_reason
        a = out.iteration_stats[-1].convergence_information[0].primal_objective
        b = single.iteration_stats[-1].convergence_information[0].primal_objective
        assert abs(a - b) <= 20 * 1e-4 * (1 + abs(b))


def test_batch_warm_start_and_limits():
    probs = mixed_problems()[:4]
    rng = np.random.default_rng(5)
    ws = [(0.01 * rng.standard_normal(p.num_row), 0.01 * rng.standard_normal(p.num_col)) for p in probs]
    ws[1] = None
    params = P.default_parameters(1e-4, iteration_limit=10)
    outs = optimize_batch(params, probs, storage=_lib.STORAGE_F64, warm_starts=ws, use_graphs=False)
    for k, (prob, out) in enumerate(zip(probs, outs)):
        ref = O.optimize(params, prob, mode=O.FUSED, warm_start=ws[k])
        assert out.iteration_count == ref["iteration_count"]
        for s_dev, s_ref in zip(out.iteration_stats, ref["stats"]):
            ci = s_dev.convergence_information[0]
            np.testing.assert_allclose(ci.dual_objective, s_ref["dual_objective"], rtol=1e-9, atol=1e-15)
            np.testing.assert_allclose(ci.l2_dual_residual, s_ref["l2_dual_residual"], rtol=1e-9, atol=1e-15)


def test_autotune_applies_the_fastest_shape():
    probs = mixed_problems()
    b = _lib.BatchSolver([p.num_row for p in probs], [p.num_col for p in probs], storage=_lib.STORAGE_F32)
    for k, p in enumerate(probs):
        b.upload(k, p.cost_matrix, p.source_distribution, p.target_distribution)
    res = autotune(b, steps=3)
    assert res["best"]["ms_per_step"] == min(t["ms_per_step"] for t in res["trials"])
    assert b.get_launch()["cols_per_tile"] == res["best"]["cols_per_tile"]
    outs = optimize_batch(P.default_parameters(1e-4, iteration_limit=400000), probs, batch=b)
    assert all(o.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL for o in outs)
    b.close()
