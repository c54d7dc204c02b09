"""Checks of the restart scheme and the primal-weight update that pin them by
hand-derived facts rather than by a second restatement.

Used by tests/test_restart_kat.py (the CPU oracle and the Python host
functions) and tests/test_gpu_restart_kat.py (the GPU host path).  A restart
trace is a list of per-evaluation records as produced by the oracle
(oracle.optimize(..., restart_cap=N)["restart_trace"]) or by the host
(pdot_amd.optimize(..., restart_trace=[])).

Facts used (src/saddle_point_gpu.jl):
  * the artificial test is avg_count >= threshold * iterations_completed
    (:568-571) and, when it fires, the adaptive test does not run, so
    kkt_reduction_ratio_last_trial is left alone;
  * the adaptive test (:492-533) restarts iff ratio < sufficient, or
    ratio < necessary and ratio > last_trial, and stores the ratio whatever
    the outcome;
  * KKT_GREEDY restarts to the average iff kkt(current) >= kkt(average)
    (:477-487);
  * update_last_restart_info! (:733-780) runs after the reset of the average
    (:667), so the distances moved are the norms of the PREVIOUS restart point:
    primal = sqrt(omega/eta) ||X_prev|| / sqrt(omega) = ||X_prev|| / sqrt(eta),
    dual = sqrt(1/(eta omega)) ||y_prev|| sqrt(omega) = ||y_prev|| / sqrt(eta),
    so with smoothing 1/2 compute_new_primal_weight (:701-728) gives
    omega' = sqrt(omega ||y_prev|| / ||X_prev||) (eta cancels), and omega' =
    omega when either distance is <= eps;
  * the restart point recorded for the next comparisons is the candidate
    (the average on a restart to the average, else the current point).
"""
from __future__ import annotations

import math

import numpy as np

NO_RESTART, TO_CURRENT, TO_AVERAGE = 1, 2, 3
EPS = float(np.finfo(np.float64).eps)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def check_trace(trace, threshold=0.36, sufficient=0.1, necessary=0.9, smoothing=0.5, exact_points=True,
                norm_rtol=1e-12):
    """Checks every record of a restart trace (ADAPTIVE_KKT + KKT_GREEDY) against
    the facts above; returns counts of the branches seen."""
    seen = {"artificial": 0, "adaptive_restart_sufficient": 0, "adaptive_restart_increase": 0,
            "adaptive_no_restart_between": 0, "adaptive_no_restart_above": 0, "to_current": 0, "to_average": 0,
            "omega_moved": 0}
    last_trial = 1.0          # create_last_restart_info (:371-393)
    prev_point = None         # pieces of the last restart point; None = the start point
    for k, t in enumerate(trace):
        assert t["last_trial_before"] == last_trial, (k, t["iteration"])
        if t["avg_count"] == 0:
            assert t["restart_used"] == NO_RESTART and t["last_trial_after"] == last_trial
            assert t["primal_weight_after"] == t["primal_weight_before"]
            continue
        artificial = t["avg_count"] >= threshold * t["iterations_completed"]
        assert bool(t["artificial"]) == artificial, (k, t["iteration"])
        # KKT_GREEDY
        to_avg = t["kkt_current"] >= t["kkt_average"]
        assert bool(t["reset_to_average"]) == to_avg, (k, t["iteration"])
        if prev_point is not None and exact_points:
            for f in ("l2_primal_residual", "l2_dual_residual", "primal_objective", "dual_objective"):
                assert t["last"][f] == prev_point[f], (k, t["iteration"], f)
        if artificial:
            seen["artificial"] += 1
            assert not t["adaptive_checked"]
            assert t["last_trial_after"] == last_trial, "last_trial must not change on an artificial restart"
            expect_restart = True
        else:
            assert t["adaptive_checked"]
            cand = t["kkt_average"] if to_avg else t["kkt_current"]
            ratio = cand / t["kkt_last"]
            assert t["ratio"] == ratio and t["last_trial_after"] == ratio, (k, t["iteration"])
            if ratio < sufficient:
                expect_restart = True
                seen["adaptive_restart_sufficient"] += 1
            elif ratio < necessary and ratio > last_trial:
                expect_restart = True
                seen["adaptive_restart_increase"] += 1
            else:
                expect_restart = False
                seen["adaptive_no_restart_between" if ratio < necessary else "adaptive_no_restart_above"] += 1
            last_trial = ratio
        if not expect_restart:
            assert t["restart_used"] == NO_RESTART, (k, t["iteration"])
            assert t["primal_weight_after"] == t["primal_weight_before"]
            continue
        assert t["restart_used"] == (TO_AVERAGE if to_avg else TO_CURRENT), (k, t["iteration"])
        seen["to_average" if to_avg else "to_current"] += 1
        # distances from the previous restart point
        if prev_point is None:
            px, py = 0.0, t["last"]["y_norm"]  # the start point: X0 = 0
            assert t["last_x_norm"] == 0.0
        else:
            px, py = prev_point["x_norm"], prev_point["y_norm"]
        assert rel(t["last_x_norm"], px) <= norm_rtol or t["last_x_norm"] == px, (k, t["iteration"])
        assert rel(t["last_y_norm"], py) <= norm_rtol or t["last_y_norm"] == py, (k, t["iteration"])
        sq = math.sqrt(t["step_size"])
        assert rel(t["primal_distance"], t["last_x_norm"] / sq) <= 1e-14 or t["primal_distance"] == 0.0
        assert rel(t["dual_distance"], t["last_y_norm"] / sq) <= 1e-14 or t["dual_distance"] == 0.0
        w0, w1 = t["primal_weight_before"], t["primal_weight_after"]
        if t["primal_distance"] > EPS and t["dual_distance"] > EPS:
            if smoothing == 0.5:
                expect = math.sqrt(w0 * t["last_y_norm"] / t["last_x_norm"])
            else:
                expect = math.exp(smoothing * math.log(t["last_y_norm"] / t["last_x_norm"])
                                  + (1 - smoothing) * math.log(w0))
            assert rel(w1, expect) <= 1e-13, (k, t["iteration"], w1, expect)
            seen["omega_moved"] += 1
        else:
            assert w1 == w0, (k, t["iteration"])
        prev_point = t["average"] if to_avg else t["current"]
    return seen


def second_iterate(C, a, b, eta=1.5):
    """The iterate after the first two accepted passes from the zero start at
    omega = 1 (pass 1 gives X = 0, u = a, v = b, r = c = 0 and eta 1 -> 1.5):
    X2 = max(0 - eta ((C - a) - b), 0), u2 = a + eta (a - 2 rowsum(X2) + 0),
    v2 = b + eta (b - 2 colsum(X2) + 0), with the reference's rounding order
    (solver_gpu.jl:125-234)."""
    X2 = np.maximum(0.0 - eta * ((C - a[:, None]) - b[None, :]), 0.0)
    r2 = X2.sum(axis=1)
    c2 = X2.sum(axis=0)
    u2 = a + eta * 1.0 * ((a - 2 * r2) + 0.0)
    v2 = b + eta * 1.0 * ((b - 2 * c2) + 0.0)
    return X2, u2, v2


def check_omega_known_answer(trace, C, a, b):
    """Restarts 1 and 2 (iterations 2 and 3, both artificial) leave omega at
    exactly 1.0: the start point and the restart-1 point (X = 0, u = a, v = b)
    have X = 0, so the primal distance is 0.  The next restart sets omega to
    sqrt(||[u2; v2]|| / ||X2||) with (X2, u2, v2) the restart-2 point, the
    iterate after two passes, in closed form (second_iterate).  Returns that
    omega."""
    by_it = {t["iteration"]: t for t in trace}
    t2, t3 = by_it[2], by_it[3]
    assert t2["restart_used"] == TO_AVERAGE and t2["artificial"]
    assert t3["restart_used"] != NO_RESTART and t3["artificial"]
    for t in (t2, t3):
        assert t["primal_weight_before"] == 1.0 and t["primal_weight_after"] == 1.0
    assert t3["last_x_norm"] == 0.0
    assert rel(t3["last_y_norm"], math.sqrt(a @ a + b @ b)) <= 1e-14
    X2, u2, v2 = second_iterate(C, a, b)
    nx = math.sqrt(float(np.sum(X2 * X2)))
    ny = math.sqrt(float(u2 @ u2 + v2 @ v2))
    later = [t for t in trace if t["iteration"] > 3 and t["restart_used"] != NO_RESTART]
    assert later, "no third restart in the trace"
    t4 = later[0]
    for t in trace:
        if t["iteration"] < t4["iteration"]:
            assert t["primal_weight_after"] == 1.0
    assert rel(t4["last_x_norm"], nx) <= 1e-13, (t4["last_x_norm"], nx)
    assert rel(t4["last_y_norm"], ny) <= 1e-13, (t4["last_y_norm"], ny)
    expect = math.sqrt(1.0 * ny / nx)
    assert rel(t4["primal_weight_after"], expect) <= 1e-13, (t4["primal_weight_after"], expect)
    return expect
