"""ctypes binding of the CPU oracle (liboracle.so, built from pdot_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline.  Product code
never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

LITERAL = 0
FUSED = 1


class OracleParams(ctypes.Structure):
    _fields_ = [
        ("primal_importance", ctypes.c_double),
        ("scale_invariant_initial_primal_weight", ctypes.c_int32),
        ("verbosity", ctypes.c_int32),
        ("record_iteration_stats", ctypes.c_int32),
        ("termination_evaluation_frequency", ctypes.c_int32),
        ("optimality_norm", ctypes.c_int32),
        ("iteration_limit", ctypes.c_int32),
        ("eps_optimal_absolute", ctypes.c_double),
        ("eps_optimal_relative", ctypes.c_double),
        ("time_sec_limit", ctypes.c_double),
        ("kkt_matrix_pass_limit", ctypes.c_double),
        ("restart_scheme", ctypes.c_int32),
        ("restart_to_current_metric", ctypes.c_int32),
        ("restart_frequency_if_fixed", ctypes.c_int64),
        ("artificial_restart_threshold", ctypes.c_double),
        ("sufficient_reduction_for_restart", ctypes.c_double),
        ("necessary_reduction_for_restart", ctypes.c_double),
        ("primal_weight_update_smoothing", ctypes.c_double),
        ("step_policy", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("reduction_exponent", ctypes.c_double),
        ("growth_exponent", ctypes.c_double),
    ]


STAT_FIELDS = [
    ("iteration_number", ctypes.c_int32),
    ("restart_used", ctypes.c_int32),
    ("candidate_type", ctypes.c_int32),
    ("cumulative_rejected_steps", ctypes.c_int32),
    ("cumulative_kkt_matrix_passes", ctypes.c_double),
    ("cumulative_time_sec", ctypes.c_double),
    ("step_size", ctypes.c_double),
    ("primal_weight", ctypes.c_double),
    ("time_spent_doing_basic_algorithm", ctypes.c_double),
    ("primal_objective", ctypes.c_double),
    ("dual_objective", ctypes.c_double),
    ("corrected_dual_objective", ctypes.c_double),
    ("l_inf_primal_residual", ctypes.c_double),
    ("l2_primal_residual", ctypes.c_double),
    ("l_inf_dual_residual", ctypes.c_double),
    ("l2_dual_residual", ctypes.c_double),
    ("relative_l_inf_primal_residual", ctypes.c_double),
    ("relative_l2_primal_residual", ctypes.c_double),
    ("relative_l_inf_dual_residual", ctypes.c_double),
    ("relative_l2_dual_residual", ctypes.c_double),
    ("relative_optimality_gap", ctypes.c_double),
    ("l_inf_primal_variable", ctypes.c_double),
    ("l2_primal_variable", ctypes.c_double),
    ("l_inf_dual_variable", ctypes.c_double),
    ("l2_dual_variable", ctypes.c_double),
]


class OracleStats(ctypes.Structure):
    _fields_ = STAT_FIELDS


RESTART_FIELDS = [
    ("iteration", ctypes.c_int32),
    ("restart_used", ctypes.c_int32),
    ("artificial", ctypes.c_int32),
    ("adaptive_checked", ctypes.c_int32),
    ("reset_to_average", ctypes.c_int32),
    ("pad", ctypes.c_int32),
    ("avg_count", ctypes.c_int64),
    ("iterations_completed", ctypes.c_int64),
    ("step_size", ctypes.c_double),
    ("primal_weight_before", ctypes.c_double),
    ("primal_weight_after", ctypes.c_double),
    ("kkt_current", ctypes.c_double),
    ("kkt_average", ctypes.c_double),
    ("kkt_last", ctypes.c_double),
    ("ratio", ctypes.c_double),
    ("last_trial_before", ctypes.c_double),
    ("last_trial_after", ctypes.c_double),
    ("last_x_norm", ctypes.c_double),
    ("last_y_norm", ctypes.c_double),
    ("primal_distance", ctypes.c_double),
    ("dual_distance", ctypes.c_double),
    ("pieces", (ctypes.c_double * 6) * 3),
]

# order of a KKT-pieces row in oracle_restart_rec.pieces
PIECE_FIELDS = ("l2_primal_residual", "l2_dual_residual", "primal_objective", "dual_objective", "x_norm", "y_norm")


class OracleRestartRec(ctypes.Structure):
    """oracle_restart_rec: what run_restart_scheme read and decided at one
    evaluation."""
    _fields_ = RESTART_FIELDS


_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_optimize.restype = ctypes.c_int
        L.oracle_optimize.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp,
                                      ctypes.c_int32, _dp, _dp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(OracleStats), ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_int32), _dp, ctypes.c_int64,
                                      ctypes.POINTER(ctypes.c_int64)]
        L.oracle_optimize_ex.restype = ctypes.c_int
        L.oracle_optimize_ex.argtypes = L.oracle_optimize.argtypes + [ctypes.POINTER(OracleRestartRec),
                                                                       ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.oracle_weighted_kkt.restype = ctypes.c_double
        L.oracle_weighted_kkt.argtypes = [ctypes.c_double] * 5
        L.oracle_should_do_adaptive_restart_kkt.restype = ctypes.c_int
        L.oracle_should_do_adaptive_restart_kkt.argtypes = [ctypes.c_double] * 4 + [_dp]
        L.oracle_restart_decision.restype = ctypes.c_int
        L.oracle_restart_decision.argtypes = [ctypes.POINTER(OracleParams), ctypes.c_int64, ctypes.c_int64,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                              ctypes.POINTER(OracleRestartRec)]
        L.oracle_restart_distances.restype = None
        L.oracle_restart_distances.argtypes = [ctypes.c_double] * 5 + [_dp, _dp]
        L.oracle_new_primal_weight.restype = ctypes.c_double
        L.oracle_new_primal_weight.argtypes = [ctypes.c_double] * 4
        L.oracle_fused_pass.restype = None
        L.oracle_fused_pass.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                        ctypes.c_int32, ctypes.c_double, ctypes.c_double, _dp, _dp, _dp, _dp, _dp,
                                        ctypes.c_double, _dp]
        L.oracle_point_reductions.restype = None
        L.oracle_point_reductions.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                              _dp, ctypes.c_int32, _dp]
        L.oracle_rounding.restype = None
        L.oracle_rounding.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp]
        L.oracle_cached_ot_info.restype = None
        L.oracle_cached_ot_info.argtypes = [ctypes.c_int64, ctypes.c_int64, _dp, _dp, _dp, _dp]
        L.oracle_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_dp)


def params_from(p) -> OracleParams:
    """Flatten a pdot_amd.PdhgParameters."""
    from pdot_amd.solver import AdaptiveStepsizeParams
    tc, rp, sp = p.termination_criteria, p.restart_params, p.step_size_policy_params
    o = OracleParams()
    o.primal_importance = p.primal_importance
    o.scale_invariant_initial_primal_weight = int(p.scale_invariant_initial_primal_weight)
    o.verbosity = int(p.verbosity)
    o.record_iteration_stats = int(p.record_iteration_stats)
    o.termination_evaluation_frequency = int(p.termination_evaluation_frequency)
    o.optimality_norm = int(tc.optimality_norm)
    o.iteration_limit = int(tc.iteration_limit)
    o.eps_optimal_absolute = tc.eps_optimal_absolute
    o.eps_optimal_relative = tc.eps_optimal_relative
    o.time_sec_limit = tc.time_sec_limit
    o.kkt_matrix_pass_limit = tc.kkt_matrix_pass_limit
    o.restart_scheme = int(rp.restart_scheme)
    o.restart_to_current_metric = int(rp.restart_to_current_metric)
    o.restart_frequency_if_fixed = int(rp.restart_frequency_if_fixed)
    o.artificial_restart_threshold = rp.artificial_restart_threshold
    o.sufficient_reduction_for_restart = rp.sufficient_reduction_for_restart
    o.necessary_reduction_for_restart = rp.necessary_reduction_for_restart
    o.primal_weight_update_smoothing = rp.primal_weight_update_smoothing
    if isinstance(sp, AdaptiveStepsizeParams):
        o.step_policy, o.reduction_exponent, o.growth_exponent = 0, float(sp.reduction_exponent), float(sp.growth_exponent)
    else:
        o.step_policy = 1
    return o


def optimize(params, problem, mode: int = FUSED, warm_start=None, stats_cap: int = 200000, trace_cap: int = 0,
             restart_cap: int = 0):
    """Run the oracle solve; returns a dict with X, u, v, termination_reason,
    iteration_count, stats (list of dicts), eta_trace and, with restart_cap
    > 0, restart_trace (list of dicts, one per non-terminating evaluation)."""
    L = lib()
    C = np.asfortranarray(problem.cost_matrix, dtype=np.float64)
    a = np.ascontiguousarray(problem.source_distribution, dtype=np.float64)
    b = np.ascontiguousarray(problem.target_distribution, dtype=np.float64)
    n, m = C.shape
    X = np.empty((n, m), dtype=np.float64, order="F")
    u = np.empty(n)
    v = np.empty(m)
    reason = ctypes.c_int32()
    iters = ctypes.c_int32()
    stats = (OracleStats * max(1, stats_cap))()
    nst = ctypes.c_int32()
    trace = np.empty(max(1, trace_cap))
    tlen = ctypes.c_int64()
    u0 = v0 = X0 = None
    if warm_start is not None:
        u0 = np.ascontiguousarray(warm_start[0], dtype=np.float64)
        v0 = np.ascontiguousarray(warm_start[1], dtype=np.float64)
        if len(warm_start) > 2 and warm_start[2] is not None:
            X0 = np.asfortranarray(warm_start[2], dtype=np.float64)
    op = params_from(params)
    rtrace = (OracleRestartRec * max(1, restart_cap))()
    nrt = ctypes.c_int32()
    rc = L.oracle_optimize_ex(ctypes.byref(op), n, m, _p(C), _p(a), _p(b), int(mode), _p(u0), _p(v0), _p(X0), _p(X),
                              _p(u), _p(v), ctypes.byref(reason), ctypes.byref(iters), stats, stats_cap,
                              ctypes.byref(nst), _p(trace), trace_cap, ctypes.byref(tlen), rtrace, restart_cap,
                              ctypes.byref(nrt))
    if rc != 0:
        raise RuntimeError(f"oracle_optimize failed with code {rc}")
    rec = [{k: getattr(stats[i], k) for k, _ in STAT_FIELDS} for i in range(min(nst.value, stats_cap))]
    rrec = [restart_rec_dict(rtrace[i]) for i in range(min(nrt.value, restart_cap))]
    return {"X": X, "u": u, "v": v, "termination_reason": reason.value, "iteration_count": iters.value,
            "stats": rec, "n_stats": nst.value, "eta_trace": trace[:min(tlen.value, trace_cap)],
            "take_step_calls": tlen.value, "restart_trace": rrec, "n_restart_trace": nrt.value}


def restart_rec_dict(r) -> dict:
    d = {k: getattr(r, k) for k, _ in RESTART_FIELDS if k not in ("pad", "pieces")}
    for q, name in enumerate(("current", "average", "last")):
        d[name] = {f: r.pieces[q][i] for i, f in enumerate(PIECE_FIELDS)}
    return d


# ---- the scalar restart functions of the oracle (saddle_point_gpu.jl:299-330, 492-780)

def weighted_kkt(omega, p2, d2, pobj, dobj) -> float:
    return float(lib().oracle_weighted_kkt(omega, p2, d2, pobj, dobj))


def should_do_adaptive_restart_kkt(candidate_kkt, last_kkt, sufficient, necessary, last_trial):
    """Returns (do_restart, last_trial after the trial)."""
    lt = ctypes.c_double(last_trial)
    r = lib().oracle_should_do_adaptive_restart_kkt(candidate_kkt, last_kkt, sufficient, necessary, ctypes.byref(lt))
    return bool(r), lt.value


def restart_decision(params, avg_count, iterations_completed, kc, ka, kl, last_trial):
    """Returns (RestartChoice value, last_trial after, record dict)."""
    op = params_from(params)
    lt = ctypes.c_double(last_trial)
    rec = OracleRestartRec()
    choice = lib().oracle_restart_decision(ctypes.byref(op), int(avg_count), int(iterations_completed), kc, ka, kl,
                                           ctypes.byref(lt), ctypes.byref(rec))
    return int(choice), lt.value, restart_rec_dict(rec)


def restart_distances(last_x_norm, last_y_norm, primal_norm_params, dual_norm_params, primal_weight):
    pd, dd = ctypes.c_double(), ctypes.c_double()
    lib().oracle_restart_distances(last_x_norm, last_y_norm, primal_norm_params, dual_norm_params, primal_weight,
                                   ctypes.byref(pd), ctypes.byref(dd))
    return pd.value, dd.value


def new_primal_weight(primal_distance, dual_distance, primal_weight, smoothing) -> float:
    return float(lib().oracle_new_primal_weight(primal_distance, dual_distance, primal_weight, smoothing))


def fused_pass(C, a, b, X, u, v, r, c, gzero, eta, omega, Xavg, uavg, vavg, ravg, cavg, S):
    """One committed pass in the fused arithmetic; arrays are updated in place.
    Returns {interaction, movement, dx2, dot, dy2}."""
    L = lib()
    n, m = C.shape
    out = np.zeros(6)
    L.oracle_fused_pass(n, m, _p(C), _p(a), _p(b), _p(X), _p(u), _p(v), _p(r), _p(c), int(gzero), float(eta),
                        float(omega), _p(Xavg), _p(uavg), _p(vavg), _p(ravg), _p(cavg), float(S), _p(out))
    return dict(zip(("interaction", "movement", "dx2", "dot", "dy2"), out[:5]))


POINT_FIELDS = ("primal_objective", "x_inf", "x_sq", "rc_inf", "rc_sq", "p_inf", "p_sq", "dual_objective",
                "u_inf", "u_sq", "y_sq")


def point_reductions(C, a, b, X, u, v, r, c, gzero=False):
    L = lib()
    n, m = C.shape
    out = np.zeros(11)
    L.oracle_point_reductions(n, m, _p(C), _p(a), _p(b), _p(X), _p(u), _p(v), _p(r), _p(c), int(gzero), _p(out))
    return dict(zip(POINT_FIELDS, out))


def rounding(a, b, X):
    L = lib()
    Xr = np.array(X, dtype=np.float64, order="F", copy=True)
    n, m = Xr.shape
    L.oracle_rounding(n, m, _p(np.ascontiguousarray(a, dtype=np.float64)),
                      _p(np.ascontiguousarray(b, dtype=np.float64)), _p(Xr))
    return Xr


def cached_ot_info(C, a, b):
    L = lib()
    C = np.asfortranarray(C, dtype=np.float64)
    out = np.zeros(4)
    L.oracle_cached_ot_info(C.shape[0], C.shape[1], _p(C), _p(np.ascontiguousarray(a, dtype=np.float64)),
                            _p(np.ascontiguousarray(b, dtype=np.float64)), _p(out))
    return out


def num_threads() -> int:
    return int(lib().oracle_num_threads())
