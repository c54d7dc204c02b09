"""The restart scheme and the primal-weight update on the GPU host path
(pdot_amd.optimize over libpdot_hip.so), pinned by hand-derived facts
(tests/restart_checks.py, src/saddle_point_gpu.jl:477-780):

* omega stays exactly 1.0 through restarts 1 and 2 and then becomes
  sqrt(||[u2; v2]|| / ||X2||) of the closed-form second iterate;
* every evaluation of whole solves follows the decision rules, and the
  distances come from the previous restart point;
* a later restart point, downloaded by an independent solve stopped at its
  evaluation, fixes the next primal weight as sqrt(omega ||[u; v]|| / ||X||);
* the host's recorded inputs fed to the oracle's scalar restart functions give
  the host's decisions bit for bit.
"""
import math

import numpy as np
import pytest

import oracle as O
import pdot_amd as P
from pdot_amd import _lib, instances
from pdot_amd.solver import define_norms

import restart_checks as RC

pytestmark = pytest.mark.gpu

NO, CUR, AVG = RC.NO_RESTART, RC.TO_CURRENT, RC.TO_AVERAGE


def _problems():
    return {"uniform_37x23": instances.uniform_sqeuclid(37, 23, seed=5),
            "gaussian_shift_60x50": instances.WORKLOADS["gaussian_shift"](60, 50, seed=3),
            "uniform_300x257": instances.uniform_sqeuclid(300, 257, seed=11)}


def _solve(prob, tol, storage=_lib.STORAGE_F64, limit=None):
    params = P.default_parameters(tol)
    if limit is not None:
        params.termination_criteria.iteration_limit = limit
    trace = []
    out = P.optimize(params, prob, storage=storage, restart_trace=trace)
    return out, trace


@pytest.mark.parametrize("name", ["uniform_37x23", "gaussian_shift_60x50", "uniform_300x257"])
@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_omega_known_answer_on_gpu(name, storage):
    prob = _problems()[name]
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    if storage == _lib.STORAGE_F32:
        C = C.astype(np.float32).astype(np.float64)  # the stored cost
    print(f"omega KAT {name} storage={storage}", flush=True)
    out, trace = _solve(prob, 1e-6, storage)
    assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
    if storage == _lib.STORAGE_F64:
        RC.check_omega_known_answer(trace, C, a, b)
    else:
        # X2 is stored in FP32: its norm is that of the rounded plan
        X2, u2, v2 = RC.second_iterate(C, a, b)
        X2 = X2.astype(np.float32).astype(np.float64)
        later = [t for t in trace if t["iteration"] > 3 and t["restart_used"] != NO][0]
        for t in trace:
            if t["iteration"] < later["iteration"]:
                assert t["primal_weight_after"] == 1.0
        nx = math.sqrt(float(np.sum(X2 * X2)))
        assert RC.rel(later["last_x_norm"], nx) <= 1e-13
        # u2 = a + 1.5 (a - 2 r2) with r2 the row sums of the stored FP32 plan
        u2 = a + 1.5 * ((a - 2 * X2.sum(axis=1)) + 0.0)
        v2 = b + 1.5 * ((b - 2 * X2.sum(axis=0)) + 0.0)
        ny = math.sqrt(float(u2 @ u2 + v2 @ v2))
        assert RC.rel(later["last_y_norm"], ny) <= 1e-13
        assert RC.rel(later["primal_weight_after"], math.sqrt(ny / nx)) <= 1e-13


@pytest.mark.parametrize("name", ["uniform_37x23", "gaussian_shift_60x50", "uniform_300x257"])
@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_gpu_trace_obeys_restart_facts(name, storage):
    prob = _problems()[name]
    print(f"trace facts {name} storage={storage}", flush=True)
    out, trace = _solve(prob, 1e-6, storage)
    assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
    assert len(trace) == len(out.iteration_stats) - 1
    assert [int(s.restart_used) for s in out.iteration_stats[:-1]] == [t["restart_used"] for t in trace]
    seen = RC.check_trace(trace)
    assert seen["artificial"] >= 3 and seen["to_average"] >= 1, seen
    assert seen["adaptive_restart_increase"] + seen["adaptive_restart_sufficient"] >= 1, seen
    assert seen["adaptive_no_restart_between"] + seen["adaptive_no_restart_above"] >= 1, seen
    assert seen["omega_moved"] >= 2, seen


@pytest.mark.parametrize("name", ["gaussian_shift_60x50", "uniform_300x257"])
def test_later_restart_point_fixes_next_primal_weight(name):
    """For each restart to the average after the second, an independent solve
    stopped by its iteration limit at that evaluation returns the restart point
    (the average); the next restart's primal weight is sqrt(omega ||[u;v]|| /
    ||X||) of those downloaded arrays."""
    prob = _problems()[name]
    out, trace = _solve(prob, 1e-6)
    restarts = [t for t in trace if t["restart_used"] != NO]
    checked = 0
    for j in range(2, len(restarts) - 1):
        rj, nxt = restarts[j], restarts[j + 1]
        if rj["restart_used"] != AVG or checked >= 3:
            continue
        k = rj["iteration"]
        print(f"{name}: restart point at iteration {k}", flush=True)
        stop, _ = _solve(prob, 1e-6, limit=k)
        assert stop.termination_reason == P.TerminationReason.TERMINATION_REASON_ITERATION_LIMIT
        assert stop.iteration_stats[-1].iteration_number == k
        X, u, v = stop.primal_solution, stop.dual_source_solution, stop.dual_target_solution
        nx = float(np.sqrt(np.sum(X * X)))
        ny = float(np.sqrt(u @ u + v @ v))
        assert RC.rel(nxt["last_x_norm"], nx) <= 1e-12, (k, nxt["last_x_norm"], nx)
        assert RC.rel(nxt["last_y_norm"], ny) <= 1e-12, (k, nxt["last_y_norm"], ny)
        if nxt["primal_distance"] > RC.EPS and nxt["dual_distance"] > RC.EPS:
            expect = math.sqrt(nxt["primal_weight_before"] * ny / nx)
            assert RC.rel(nxt["primal_weight_after"], expect) <= 1e-12
        checked += 1
    assert checked >= 1


@pytest.mark.parametrize("name", ["gaussian_shift_60x50", "uniform_300x257"])
def test_gpu_trace_replayed_through_oracle_functions(name):
    """The host's recorded restart inputs, fed to the oracle's scalar restart
    functions (the code the oracle's own solve runs), give the host's weighted
    KKTs, decisions, last trials, distances and primal weights bit for bit."""
    prob = _problems()[name]
    params = P.default_parameters(1e-6)
    out, trace = _solve(prob, 1e-6)
    assert len(trace) > 20
    for t in trace:
        w = t["primal_weight_before"]
        kk = {}
        for q in ("current", "average", "last"):
            p = t[q]
            kk[q] = O.weighted_kkt(w, p["l2_primal_residual"], p["l2_dual_residual"], p["primal_objective"],
                                   p["dual_objective"])
        choice, lt, rec = O.restart_decision(params, t["avg_count"], t["iterations_completed"], kk["current"],
                                             kk["average"], kk["last"], t["last_trial_before"])
        assert choice == t["restart_used"], t["iteration"]
        assert lt == t["last_trial_after"], t["iteration"]
        if t["avg_count"] == 0:
            continue
        assert kk["current"] == t["kkt_current"] and kk["average"] == t["kkt_average"]
        if rec["adaptive_checked"]:
            assert kk["last"] == t["kkt_last"] and rec["ratio"] == t["ratio"]
        if choice != NO:
            pnp, dnp = define_norms(t["step_size"], w)
            pd, dd = O.restart_distances(t["last_x_norm"], t["last_y_norm"], pnp, dnp, w)
            assert (pd, dd) == (t["primal_distance"], t["dual_distance"]), t["iteration"]
            assert O.new_primal_weight(pd, dd, w, 0.5) == t["primal_weight_after"], t["iteration"]
