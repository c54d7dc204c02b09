"""Restart scheme and primal-weight update: the host-scalar half of
src/saddle_point_gpu.jl.  The n*m work behind it (weighted KKT pieces, the
restart to the average, the reset of the average) runs in libpdot_hip.so;
this module keeps the decisions.

The last restart point is held as scalars (||X||, ||[u;v]|| and its KKT
pieces): the reference recomputes the weighted KKT of that point at every
evaluation with only omega changing (saddle_point_gpu.jl:503-515), and the
distances it feeds to the primal-weight update are norms of the previous
restart point because the average is reset before update_last_restart_info!
reads it (:667-688, 753-767).
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .solve_log import RestartChoice


class RestartScheme(enum.IntEnum):
    """saddle_point_gpu.jl:401"""
    NO_RESTARTS = 0
    FIXED_FREQUENCY = 1
    ADAPTIVE_KKT = 2


class RestartToCurrentMetric(enum.IntEnum):
    """saddle_point_gpu.jl:408"""
    NO_RESTART_TO_CURRENT = 0
    KKT_GREEDY = 1


NO_RESTARTS = RestartScheme.NO_RESTARTS
FIXED_FREQUENCY = RestartScheme.FIXED_FREQUENCY
ADAPTIVE_KKT = RestartScheme.ADAPTIVE_KKT
NO_RESTART_TO_CURRENT = RestartToCurrentMetric.NO_RESTART_TO_CURRENT
KKT_GREEDY = RestartToCurrentMetric.KKT_GREEDY

EPS = float(np.finfo(np.float64).eps)


@dataclass
class RestartParameters:
    """saddle_point_gpu.jl:411-441"""
    restart_scheme: RestartScheme
    restart_to_current_metric: RestartToCurrentMetric
    restart_frequency_if_fixed: int
    artificial_restart_threshold: float
    sufficient_reduction_for_restart: float
    necessary_reduction_for_restart: float
    primal_weight_update_smoothing: float


def construct_restart_parameters(restart_scheme, restart_to_current_metric, restart_frequency_if_fixed,
                                 artificial_restart_threshold, sufficient_reduction_for_restart,
                                 necessary_reduction_for_restart, primal_weight_update_smoothing) -> RestartParameters:
    """saddle_point_gpu.jl:446-472, with its assertions."""
    assert restart_frequency_if_fixed > 1
    assert 0.0 < artificial_restart_threshold <= 1.0
    assert 0.0 < sufficient_reduction_for_restart <= necessary_reduction_for_restart <= 1.0
    assert 0.0 <= primal_weight_update_smoothing <= 1.0
    return RestartParameters(RestartScheme(restart_scheme), RestartToCurrentMetric(restart_to_current_metric),
                             int(restart_frequency_if_fixed), float(artificial_restart_threshold),
                             float(sufficient_reduction_for_restart), float(necessary_reduction_for_restart),
                             float(primal_weight_update_smoothing))


@dataclass
class KKTPieces:
    """The pieces of compute_weight_kkt_residual (saddle_point_gpu.jl:299-330)
    of one point, plus the norms its restart distances need."""
    l2_primal_residual: float
    l2_dual_residual: float
    primal_objective: float
    dual_objective: float
    x_norm: float
    y_norm: float

    @staticmethod
    def from_point_stats(ps) -> "KKTPieces":
        return KKTPieces(math.sqrt(ps.p_sq), math.sqrt(ps.rc_sq), ps.primal_objective, ps.dual_objective,
                         math.sqrt(ps.x_sq), math.sqrt(ps.y_sq))

    def weighted_kkt(self, primal_weight: float) -> float:
        """sqrt(w P^2 + D^2/w + |pobj - dobj|^2), saddle_point_gpu.jl:327"""
        p2, d2 = self.l2_primal_residual, self.l2_dual_residual
        g = abs(self.primal_objective - self.dual_objective)
        return math.sqrt(primal_weight * (p2 * p2) + 1 / primal_weight * (d2 * d2) + g * g)


@dataclass
class RestartInfo:
    """CuRestartInfo (saddle_point_gpu.jl:332-366), kept as scalars."""
    point: KKTPieces
    last_restart_length: int = 1
    primal_distance_moved_last_restart_period: float = 0.0
    dual_distance_moved_last_restart_period: float = 0.0
    kkt_reduction_ratio_last_trial: float = 1.0


def create_last_restart_info(start: KKTPieces) -> RestartInfo:
    """saddle_point_gpu.jl:371-393"""
    return RestartInfo(point=start)


def should_reset_to_average(current_kkt: float, average_kkt: float, metric: RestartToCurrentMetric) -> bool:
    """saddle_point_gpu.jl:477-487"""
    if metric == KKT_GREEDY:
        return current_kkt >= average_kkt
    return True


def should_do_adaptive_restart_kkt(candidate_kkt: float, restart_params: RestartParameters,
                                   last_restart_info: RestartInfo, primal_weight: float) -> bool:
    """saddle_point_gpu.jl:492-533"""
    kkt_last = last_restart_info.point.weighted_kkt(primal_weight)
    do_restart = False
    ratio = candidate_kkt / kkt_last if kkt_last != 0.0 else (math.inf if candidate_kkt > 0 else math.nan)
    if ratio < restart_params.necessary_reduction_for_restart:
        if ratio < restart_params.sufficient_reduction_for_restart:
            do_restart = True
        elif ratio > last_restart_info.kkt_reduction_ratio_last_trial:
            do_restart = True
    last_restart_info.kkt_reduction_ratio_last_trial = ratio
    return do_restart


@dataclass
class RestartDecision:
    """The outcome of run_restart_scheme and what it read: the weighted KKT
    residuals (NaN where not formed), the ratio of the adaptive test (NaN
    unless it ran), kkt_reduction_ratio_last_trial before and after, and on a
    restart the norms of the previous restart point and the distances moved."""
    choice: RestartChoice
    to_average: bool = False
    artificial: bool = False
    adaptive_checked: bool = False
    reset_to_average: bool = False
    kkt_current: float = math.nan
    kkt_average: float = math.nan
    kkt_last: float = math.nan
    ratio: float = math.nan
    last_trial_before: float = math.nan
    last_trial_after: float = math.nan
    last_x_norm: float = math.nan
    last_y_norm: float = math.nan
    primal_distance: float = math.nan
    dual_distance: float = math.nan


def run_restart_scheme(avg_count: int, current: Optional[KKTPieces], average: Optional[KKTPieces],
                       last_restart_info: RestartInfo, iterations_completed: int, primal_norm_params: float,
                       dual_norm_params: float, primal_weight: float,
                       restart_params: RestartParameters) -> RestartDecision:
    """saddle_point_gpu.jl:539-696.  `current`/`average` are the KKT pieces of
    the current point and of the average, evaluated on the device.  On a
    restart the caller moves the average into the current point (if
    to_average) and resets the average; this function updates the restart
    info as update_last_restart_info! (:733-780) does."""
    out = RestartDecision(RestartChoice.RESTART_CHOICE_NO_RESTART,
                          last_trial_before=last_restart_info.kkt_reduction_ratio_last_trial)
    if not avg_count > 0:
        out.last_trial_after = last_restart_info.kkt_reduction_ratio_last_trial
        return out
    restart_length = avg_count
    do_restart = False
    if restart_length >= restart_params.artificial_restart_threshold * iterations_completed:
        do_restart = True
        out.artificial = True
    if restart_params.restart_scheme == NO_RESTARTS:
        reset_to_average = False
        candidate = None
    else:
        kc = current.weighted_kkt(primal_weight)
        ka = average.weighted_kkt(primal_weight)
        out.kkt_current, out.kkt_average = kc, ka
        reset_to_average = should_reset_to_average(kc, ka, restart_params.restart_to_current_metric)
        candidate = (average, ka) if reset_to_average else (current, kc)
    out.reset_to_average = reset_to_average
    if not do_restart:
        if restart_params.restart_scheme == ADAPTIVE_KKT:
            out.kkt_last = last_restart_info.point.weighted_kkt(primal_weight)
            do_restart = should_do_adaptive_restart_kkt(candidate[1], restart_params, last_restart_info,
                                                        primal_weight)
            out.adaptive_checked = True
            out.ratio = last_restart_info.kkt_reduction_ratio_last_trial
        elif (restart_params.restart_scheme == FIXED_FREQUENCY
              and restart_params.restart_frequency_if_fixed <= restart_length):
            do_restart = True
    out.last_trial_after = last_restart_info.kkt_reduction_ratio_last_trial
    if not do_restart:
        return out
    new_point = average if reset_to_average else current
    # update_last_restart_info!: avg - last = -last after the reset (:753-767)
    lp = last_restart_info.point
    out.last_x_norm, out.last_y_norm = lp.x_norm, lp.y_norm
    last_restart_info.primal_distance_moved_last_restart_period = (
        math.sqrt(primal_norm_params) * lp.x_norm / math.sqrt(primal_weight))
    last_restart_info.dual_distance_moved_last_restart_period = (
        math.sqrt(dual_norm_params) * lp.y_norm * math.sqrt(primal_weight))
    out.primal_distance = last_restart_info.primal_distance_moved_last_restart_period
    out.dual_distance = last_restart_info.dual_distance_moved_last_restart_period
    last_restart_info.point = new_point
    last_restart_info.last_restart_length = restart_length
    if reset_to_average:
        out.choice, out.to_average = RestartChoice.RESTART_CHOICE_RESTART_TO_AVERAGE, True
    else:
        out.choice = RestartChoice.RESTART_CHOICE_WEIGHTED_AVERAGE_RESET
    return out


def compute_new_primal_weight(last_restart_info: RestartInfo, primal_weight: float,
                              primal_weight_update_smoothing: float, verbosity: int = 0) -> float:
    """saddle_point_gpu.jl:701-728"""
    primal_distance = last_restart_info.primal_distance_moved_last_restart_period
    dual_distance = last_restart_info.dual_distance_moved_last_restart_period
    if primal_distance > EPS and dual_distance > EPS:
        new_primal_weight_estimate = dual_distance / primal_distance
        log_primal_weight = (primal_weight_update_smoothing * math.log(new_primal_weight_estimate)
                             + (1 - primal_weight_update_smoothing) * math.log(primal_weight))
        primal_weight = math.exp(log_primal_weight)
        if verbosity >= 4:
            print(f"  New computed primal weight is {primal_weight:.2e}")
        return primal_weight
    return primal_weight


def select_initial_primal_weight(cost_fro_norm: float, rhs_l2_norm: float, primal_norm_params: float,
                                 dual_norm_params: float, primal_importance: float, verbosity: int = 0) -> float:
    """saddle_point_gpu.jl:848-866"""
    rhs_vec_norm = math.sqrt(dual_norm_params) * rhs_l2_norm
    obj_vec_norm = math.sqrt(primal_norm_params) * cost_fro_norm
    if obj_vec_norm > 0.0 and rhs_vec_norm > 0.0:
        primal_weight = primal_importance * (obj_vec_norm / rhs_vec_norm)
    else:
        primal_weight = primal_importance
    if verbosity >= 6:
        print(f"Initial primal weight = {primal_weight}")
    return primal_weight
