"""Known answers for the restart scheme and the primal-weight update
(src/saddle_point_gpu.jl:477-780), worked by hand, run on the oracle's
exported scalar functions (the code its solve runs) and on the Python host
functions that drive the GPU (pdot_amd/saddle_point.py); then whole oracle
solves checked against the same facts (tests/restart_checks.py) and the
oracle's per-evaluation restart trace replayed through the host functions.
"""
import math

import numpy as np
import pytest

import oracle as O
import pdot_amd as P
from pdot_amd import instances
from pdot_amd.saddle_point import (ADAPTIVE_KKT, FIXED_FREQUENCY, KKT_GREEDY, NO_RESTART_TO_CURRENT, NO_RESTARTS,
                                   KKTPieces, RestartInfo, compute_new_primal_weight, construct_restart_parameters,
                                   run_restart_scheme, should_do_adaptive_restart_kkt)
from pdot_amd.solver import define_norms

import restart_checks as RC

NO, CUR, AVG = RC.NO_RESTART, RC.TO_CURRENT, RC.TO_AVERAGE


def params_with(scheme=ADAPTIVE_KKT, metric=KKT_GREEDY, freq=1000):
    p = P.default_parameters(1e-4)
    p.restart_params = construct_restart_parameters(scheme, metric, freq, 0.36, 0.1, 0.9, 0.5)
    return p


def pieces(k, x_norm=1.0, y_norm=1.0):
    """KKT pieces whose weighted KKT is k at every primal weight: no residuals,
    objective gap k."""
    return KKTPieces(0.0, 0.0, k, 0.0, x_norm, y_norm)


# (name, scheme, metric, freq, avg_count, iterations_completed, kc, ka, kl, last_trial_in,
#  expected choice, expected last_trial_out).  Thresholds are the reference's
#  defaults (scripts/solve.jl:58-66): artificial 0.36, sufficient 0.1, necessary 0.9.
TABLE = [
    # not artificial (10 < 0.36 * 100): the adaptive ratio rule decides
    ("ratio below sufficient", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 0.5, 0.04, 1.0, 1.0, AVG, 0.04),
    ("between, above last trial", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 0.5, 0.5, 1.0, 0.25, AVG, 0.5),
    ("between, below last trial", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 0.5, 0.5, 1.0, 0.75, NO, 0.5),
    ("between, equal to last trial", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 0.5, 0.5, 1.0, 0.5, NO, 0.5),
    ("between, first trial (last trial 1.0)", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 0.5, 0.5, 1.0, 1.0, NO, 0.5),
    ("at necessary", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 1.0, 0.9, 1.0, 0.25, NO, 0.9),
    ("above one", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 4.0, 2.0, 1.0, 0.25, NO, 2.0),
    ("current better: candidate is the current point", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100,
     0.0625, 0.5, 1.0, 1.0, CUR, 0.0625),
    ("current better, between, above last trial", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100,
     0.5, 0.75, 1.0, 0.25, CUR, 0.5),
    ("ties go to the average", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 0.05, 0.05, 1.0, 1.0, AVG, 0.05),
    ("last point scales the ratio", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 1.0, 0.5, 8.0, 1.0, AVG, 0.0625),
    ("NO_RESTART_TO_CURRENT ignores a better current point", ADAPTIVE_KKT, NO_RESTART_TO_CURRENT, 1000, 10, 100,
     0.0625, 0.5, 1.0, 0.25, AVG, 0.5),
    ("NO_RESTART_TO_CURRENT, ratio of the average above necessary", ADAPTIVE_KKT, NO_RESTART_TO_CURRENT, 1000,
     10, 100, 0.0625, 1.0, 1.0, 0.25, NO, 1.0),
    ("last point at the optimum: ratio inf", ADAPTIVE_KKT, KKT_GREEDY, 1000, 10, 100, 1.0, 0.5, 0.0, 0.25, NO,
     math.inf),
    # artificial (avg_count >= 0.36 * iterations_completed): restart, last trial untouched
    ("artificial at the boundary", ADAPTIVE_KKT, KKT_GREEDY, 1000, 36, 100, 4.0, 2.0, 1.0, 0.75, AVG, 0.75),
    ("just below the artificial boundary", ADAPTIVE_KKT, KKT_GREEDY, 1000, 35, 100, 4.0, 2.0, 1.0, 0.75, NO, 2.0),
    ("artificial to the current point", ADAPTIVE_KKT, KKT_GREEDY, 1000, 50, 100, 0.125, 0.25, 1.0, 0.75, CUR, 0.75),
    ("artificial on the first evaluation after the start", ADAPTIVE_KKT, KKT_GREEDY, 1000, 1, 1, 2.0, 2.0, 1.0,
     1.0, AVG, 1.0),
    ("empty average", ADAPTIVE_KKT, KKT_GREEDY, 1000, 0, 0, 0.01, 0.01, 1.0, 0.75, NO, 0.75),
    # NO_RESTARTS: only the artificial restarts, always to the current point
    ("NO_RESTARTS, not artificial", NO_RESTARTS, KKT_GREEDY, 1000, 10, 100, 1.0, 0.01, 1.0, 0.75, NO, 0.75),
    ("NO_RESTARTS, artificial", NO_RESTARTS, KKT_GREEDY, 1000, 50, 100, 1.0, 0.01, 1.0, 0.75, CUR, 0.75),
    # FIXED_FREQUENCY: restart once the average holds `freq` iterates
    ("FIXED_FREQUENCY at the frequency", FIXED_FREQUENCY, KKT_GREEDY, 10, 10, 100, 1.0, 0.5, 1.0, 0.75, AVG, 0.75),
    ("FIXED_FREQUENCY below the frequency", FIXED_FREQUENCY, KKT_GREEDY, 10, 9, 100, 1.0, 0.5, 1.0, 0.75, NO, 0.75),
    ("FIXED_FREQUENCY to the current point", FIXED_FREQUENCY, KKT_GREEDY, 10, 12, 100, 0.25, 0.5, 1.0, 0.75, CUR,
     0.75),
]


@pytest.mark.parametrize("row", TABLE, ids=[r[0] for r in TABLE])
def test_restart_decision_table_oracle(row):
    name, scheme, metric, freq, cnt, iters, kc, ka, kl, lt_in, choice, lt_out = row
    got, lt, rec = O.restart_decision(params_with(scheme, metric, freq), cnt, iters, kc, ka, kl, lt_in)
    assert got == choice, name
    assert lt == lt_out, name
    assert rec["artificial"] == int(cnt > 0 and cnt >= 0.36 * iters)


@pytest.mark.parametrize("row", TABLE, ids=[r[0] for r in TABLE])
def test_restart_decision_table_host(row):
    name, scheme, metric, freq, cnt, iters, kc, ka, kl, lt_in, choice, lt_out = row
    for k in (kc, ka, kl):
        assert pieces(k).weighted_kkt(3.0) == k  # sqrt(k*k) == k
    cur, avg = pieces(kc, 2.0, 3.0), pieces(ka, 5.0, 7.0)
    info = RestartInfo(point=pieces(kl, 0.5, 0.25), kkt_reduction_ratio_last_trial=lt_in)
    pnp, dnp = define_norms(4.0, 1.0)
    d = run_restart_scheme(cnt, cur, avg if cnt > 0 else None, info, iters, pnp, dnp, 1.0,
                           params_with(scheme, metric, freq).restart_params)
    assert int(d.choice) == choice, name
    assert info.kkt_reduction_ratio_last_trial == lt_out, name
    if choice == NO:
        assert info.point.primal_objective == kl  # the last restart point is kept
    else:
        # the new restart point is the candidate; the distances are those of the
        # previous one: ||X|| / sqrt(eta) = 0.5 / 2, ||y|| / sqrt(eta) = 0.25 / 2
        assert info.point is (avg if choice == AVG else cur), name
        assert info.primal_distance_moved_last_restart_period == 0.25
        assert info.dual_distance_moved_last_restart_period == 0.125
        assert info.last_restart_length == cnt


@pytest.mark.parametrize("cand,last,lt_in,expect,lt_out", [
    (0.05, 1.0, 1.0, True, 0.05), (0.5, 1.0, 0.25, True, 0.5), (0.5, 1.0, 0.75, False, 0.5),
    (0.95, 1.0, 0.25, False, 0.95), (3.0, 1.0, 0.25, False, 3.0), (0.45, 4.5, 1.0, False, 0.1),
    (0.225, 4.5, 1.0, True, 0.05)])
def test_adaptive_rule_alone(cand, last, lt_in, expect, lt_out):
    """should_do_adaptive_restart_kkt (:492-533) with the reference's thresholds."""
    got, lt = O.should_do_adaptive_restart_kkt(cand, last, 0.1, 0.9, lt_in)
    assert (got, lt) == (expect, lt_out)
    info = RestartInfo(point=pieces(last), kkt_reduction_ratio_last_trial=lt_in)
    assert should_do_adaptive_restart_kkt(cand, params_with().restart_params, info, 2.0) == expect
    assert info.kkt_reduction_ratio_last_trial == lt_out


def test_weighted_kkt_known_answers():
    """compute_weight_kkt_residual (:299-330): sqrt(w P^2 + D^2 / w + gap^2)."""
    cases = [((4.0, 1.0, 2.0, 3.0, 1.0), 3.0),          # 4*1 + 4/4 + 4 = 9
             ((0.25, 2.0, 1.0, 0.0, 0.0), math.sqrt(5.0)),  # 0.25*4 + 1/0.25 = 5
             ((1.0, 0.0, 0.0, -1.5, 0.5), 2.0)]
    for args, expect in cases:
        assert O.weighted_kkt(*args) == expect
        w, p2, d2, po, do = args
        assert KKTPieces(p2, d2, po, do, 0.0, 0.0).weighted_kkt(w) == expect


def test_distances_and_new_primal_weight_known_answers():
    """update_last_restart_info! distances (:753-767) and compute_new_primal_weight
    (:701-728), worked by hand: eta = 1, omega = 4 gives primal_norm_params 4 and
    dual_norm_params 1/4; ||X_prev|| = 3, ||y_prev|| = 36 give distances
    2 * 3 / 2 = 3 and 0.5 * 36 * 2 = 36, estimate 12, and with smoothing 1/2
    omega' = sqrt(12 * 4) = sqrt(48)."""
    pnp, dnp = define_norms(1.0, 4.0)
    assert (pnp, dnp) == (4.0, 0.25)
    pd, dd = O.restart_distances(3.0, 36.0, pnp, dnp, 4.0)
    assert (pd, dd) == (3.0, 36.0)
    w = O.new_primal_weight(pd, dd, 4.0, 0.5)
    assert w == pytest.approx(math.sqrt(48.0), rel=1e-15)
    info = RestartInfo(point=pieces(1.0), primal_distance_moved_last_restart_period=pd,
                       dual_distance_moved_last_restart_period=dd)
    assert compute_new_primal_weight(info, 4.0, 0.5) == w
    # eta = 4: the distances shrink by sqrt(eta) = 2; omega' does not change
    pd4, dd4 = O.restart_distances(3.0, 36.0, *define_norms(4.0, 4.0), 4.0)
    assert (pd4, dd4) == (1.5, 18.0)
    assert O.new_primal_weight(pd4, dd4, 4.0, 0.5) == pytest.approx(math.sqrt(48.0), rel=1e-15)
    # smoothing 1: the estimate itself; smoothing 0: frozen
    assert O.new_primal_weight(pd, dd, 4.0, 1.0) == pytest.approx(12.0, rel=1e-15)
    assert O.new_primal_weight(pd, dd, 4.0, 0.0) == pytest.approx(4.0, rel=1e-15)
    # a distance at or below eps(Float64) leaves omega unchanged
    for d in [(0.0, 36.0), (3.0, 0.0), (RC.EPS, 36.0), (3.0, RC.EPS)]:
        assert O.new_primal_weight(d[0], d[1], 4.0, 0.5) == 4.0
        i2 = RestartInfo(point=pieces(1.0), primal_distance_moved_last_restart_period=d[0],
                         dual_distance_moved_last_restart_period=d[1])
        assert compute_new_primal_weight(i2, 4.0, 0.5) == 4.0
    assert O.new_primal_weight(2 * RC.EPS, 36.0, 4.0, 0.5) != 4.0


def _instances():
    return [instances.uniform_sqeuclid(37, 23, seed=5), instances.WORKLOADS["gaussian_shift"](60, 50, seed=3)]


def _prob(p):
    return p.problem if hasattr(p, "problem") else p


@pytest.mark.parametrize("mode", [O.LITERAL, O.FUSED])
@pytest.mark.parametrize("which", [0, 1])
def test_oracle_trace_obeys_restart_facts(mode, which):
    """A whole solve to 1e-6: every evaluation's restart decision, last trial,
    restart point and primal weight follow the hand-derived rules, and the trace
    holds a restart to the average and to the current point, non-artificial
    restarts by both adaptive branches, and both kinds of non-restart."""
    prob = _prob(_instances()[which])
    r = O.optimize(P.default_parameters(1e-6), prob, mode=mode, restart_cap=100000)
    assert r["termination_reason"] == int(P.TerminationReason.TERMINATION_REASON_OPTIMAL)
    trace = r["restart_trace"]
    assert len(trace) == r["n_restart_trace"] == len(r["stats"]) - 1
    seen = RC.check_trace(trace)
    assert seen["artificial"] >= 3 and seen["to_average"] >= 1 and seen["to_current"] >= 1, seen
    assert seen["adaptive_restart_increase"] + seen["adaptive_restart_sufficient"] >= 1, seen
    assert seen["adaptive_no_restart_between"] >= 1 and seen["adaptive_no_restart_above"] >= 1, seen
    assert seen["omega_moved"] >= 3, seen
    # restart_used in the IterationStats agrees with the trace
    assert [s["restart_used"] for s in r["stats"][:-1]] == [t["restart_used"] for t in trace]


@pytest.mark.parametrize("mode", [O.LITERAL, O.FUSED])
@pytest.mark.parametrize("which", [0, 1])
def test_oracle_omega_known_answer(mode, which):
    """omega stays exactly 1.0 through restarts 1 and 2, then becomes
    sqrt(||[u2; v2]|| / ||X2||) of the closed-form second iterate."""
    prob = _prob(_instances()[which])
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    r = O.optimize(P.default_parameters(1e-6), prob, mode=mode, restart_cap=100000)
    # a rejected pass is replayed with the unchanged step size (solver_gpu.jl:300-306),
    # so pass 2 is taken at eta = 1.5 whether or not it is first rejected
    w = RC.check_omega_known_answer(r["restart_trace"], C, a, b)
    assert w != 1.0


def _host_replay(trace, params):
    """Feeds each recorded evaluation of an oracle trace to the host's
    run_restart_scheme / compute_new_primal_weight; returns the mismatches."""
    rp = params.restart_params
    bad = []
    for t in trace:
        if t["avg_count"] == 0:
            continue
        cur = KKTPieces(*(t["current"][f] for f in O.PIECE_FIELDS))
        avg = KKTPieces(*(t["average"][f] for f in O.PIECE_FIELDS))
        last = KKTPieces(*(t["last"][f] for f in O.PIECE_FIELDS))
        if t["restart_used"] != NO:
            last.x_norm, last.y_norm = t["last_x_norm"], t["last_y_norm"]
        info = RestartInfo(point=last, kkt_reduction_ratio_last_trial=t["last_trial_before"])
        pnp, dnp = define_norms(t["step_size"], t["primal_weight_before"])
        d = run_restart_scheme(t["avg_count"], cur, avg, info, t["iterations_completed"], pnp, dnp,
                               t["primal_weight_before"], rp)
        w = t["primal_weight_before"]
        if d.choice != P.RestartChoice.RESTART_CHOICE_NO_RESTART:
            w = compute_new_primal_weight(info, w, rp.primal_weight_update_smoothing)
        got = (int(d.choice), info.kkt_reduction_ratio_last_trial, w, d.kkt_current, d.kkt_average)
        want = (t["restart_used"], t["last_trial_after"], t["primal_weight_after"], t["kkt_current"],
                t["kkt_average"])
        if not all(x == y or (x != x and y != y) for x, y in zip(got, want)):
            bad.append((t["iteration"], got, want))
        if d.choice != P.RestartChoice.RESTART_CHOICE_NO_RESTART:
            if (d.primal_distance, d.dual_distance) != (t["primal_distance"], t["dual_distance"]):
                bad.append((t["iteration"], "distances", d.primal_distance, t["primal_distance"]))
    return bad


@pytest.mark.parametrize("mode", [O.LITERAL, O.FUSED])
def test_oracle_trace_replayed_through_host_functions(mode):
    """The host's restart functions make the oracle's decisions, bit for bit,
    at every evaluation of whole solves under several restart policies."""
    prob = _prob(_instances()[1])
    for scheme, metric in [(ADAPTIVE_KKT, KKT_GREEDY), (ADAPTIVE_KKT, NO_RESTART_TO_CURRENT),
                           (FIXED_FREQUENCY, KKT_GREEDY), (NO_RESTARTS, KKT_GREEDY)]:
        params = params_with(scheme, metric, 64)
        params.termination_criteria.eps_optimal_absolute = 1e-6
        params.termination_criteria.eps_optimal_relative = 1e-6
        params.termination_criteria.iteration_limit = 20000
        r = O.optimize(params, prob, mode=mode, restart_cap=100000)
        trace = r["restart_trace"]
        assert len(trace) > 10
        assert _host_replay(trace, params) == [], (scheme, metric)


def test_planted_mistakes_fail_the_checks():
    """The trace checks reject the plausible slips: primal and dual distances
    swapped, the new restart point used for the distances, last_trial updated
    on an artificial restart."""
    prob = _prob(_instances()[0])
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    trace = O.optimize(P.default_parameters(1e-6), prob, mode=O.FUSED, restart_cap=100000)["restart_trace"]
    RC.check_trace(trace)
    moved = [i for i, t in enumerate(trace) if t["restart_used"] != NO and t["primal_distance"] > RC.EPS]
    # swapped distances: omega' = sqrt(omega ||X|| / ||y||)
    bad = [dict(t) for t in trace]
    k = moved[0]
    t = bad[k]
    t["primal_weight_after"] = math.sqrt(t["primal_weight_before"] * t["last_x_norm"] / t["last_y_norm"])
    with pytest.raises(AssertionError):
        RC.check_trace(bad)
    # the new restart point's norms instead of the previous one's
    bad = [dict(t) for t in trace]
    t = bad[k]
    new_pt = t["average"] if t["restart_used"] == AVG else t["current"]
    t["last_x_norm"], t["last_y_norm"] = new_pt["x_norm"], new_pt["y_norm"]
    with pytest.raises(AssertionError):
        RC.check_trace(bad)
    # last_trial updated on an artificial restart
    bad = [dict(t) for t in trace]
    art = [i for i, t in enumerate(bad) if t["artificial"] and i + 1 < len(bad)]
    i = art[-1]
    bad[i]["last_trial_after"] = 0.5
    for t in bad[i + 1:]:
        t["last_trial_before"] = 0.5
        break
    with pytest.raises(AssertionError):
        RC.check_trace(bad)
    # the omega known answer rejects a wrong second iterate (pass 2 at eta = 1)
    with pytest.raises(AssertionError):
        X2, u2, v2 = RC.second_iterate(C, a, b, eta=1.0)
        nx = math.sqrt(float(np.sum(X2 * X2)))
        ny = math.sqrt(float(u2 @ u2 + v2 @ v2))
        later = [t for t in trace if t["iteration"] > 3 and t["restart_used"] != NO][0]
        assert RC.rel(later["primal_weight_after"], math.sqrt(ny / nx)) <= 1e-13
