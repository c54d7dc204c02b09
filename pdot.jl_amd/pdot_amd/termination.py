"""Termination criteria, restating src/termination.jl."""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .solve_log import IterationStats, TerminationReason, ConvergenceInformation

INT32_MAX = 2**31 - 1


class OptimalityNorm(enum.IntEnum):
    """termination.jl:2"""
    L_INF = 0
    L_2 = 1


L_INF = OptimalityNorm.L_INF
L_2 = OptimalityNorm.L_2


@dataclass
class TerminationCriteria:
    """termination.jl:5-51"""
    optimality_norm: OptimalityNorm
    eps_optimal_absolute: float
    eps_optimal_relative: float
    time_sec_limit: float
    iteration_limit: int
    kkt_matrix_pass_limit: float


def construct_termination_criteria(*, optimality_norm=L_2, eps_optimal_absolute, eps_optimal_relative,
                                   time_sec_limit=math.inf, iteration_limit=INT32_MAX,
                                   kkt_matrix_pass_limit=math.inf) -> TerminationCriteria:
    """termination.jl:53-69 (keyword constructor with the same defaults)."""
    return TerminationCriteria(OptimalityNorm(optimality_norm), float(eps_optimal_absolute),
                               float(eps_optimal_relative), float(time_sec_limit), int(iteration_limit),
                               float(kkt_matrix_pass_limit))


def validate_termination_criteria(criteria: TerminationCriteria) -> None:
    """termination.jl:71-81"""
    if criteria.time_sec_limit <= 0:
        raise ValueError("time_sec_limit must be positive")
    if criteria.iteration_limit <= 0:
        raise ValueError("iteration_limit must be positive")
    if criteria.kkt_matrix_pass_limit <= 0:
        raise ValueError("kkt_matrix_pass_limit must be positive")


@dataclass(frozen=True)
class CachedOTInfo:
    """termination.jl:87-92"""
    l_inf_norm_primal_linear_objective: float
    l_inf_norm_primal_right_hand_side: float
    l2_norm_primal_linear_objective: float
    l2_norm_primal_right_hand_side: float


def cached_ot_info(problem) -> CachedOTInfo:
    """termination.jl:94-101: ||C||_inf, ||[a;b]||_inf, ||C||_F, ||[a;b]||_2
    (Julia's norm(matrix, p) treats the matrix as a vector)."""
    C = problem.cost_matrix
    ab = np.concatenate([np.asarray(problem.source_distribution, dtype=np.float64),
                         np.asarray(problem.target_distribution, dtype=np.float64)])
    c_inf = float(np.max(np.abs(C))) if C.size else 0.0
    c_fro = float(np.sqrt(np.einsum("ij,ij->", C, C, optimize=False))) if C.size else 0.0
    return CachedOTInfo(c_inf, float(np.max(np.abs(ab))) if ab.size else 0.0, c_fro, float(np.sqrt(ab @ ab)))


def optimality_criteria_met(optimality_norm: OptimalityNorm, abs_tol: float, rel_tol: float,
                            convergence_information: ConvergenceInformation, ot_cache: CachedOTInfo) -> bool:
    """termination.jl:106-134 (strict inequalities)."""
    ci = convergence_information
    abs_obj = abs(ci.primal_objective) + abs(ci.dual_objective)
    gap = abs(ci.primal_objective - ci.dual_objective)
    if optimality_norm == L_INF:
        primal_err, primal_base = ci.l_inf_primal_residual, ot_cache.l_inf_norm_primal_right_hand_side
        dual_err, dual_base = ci.l_inf_dual_residual, ot_cache.l_inf_norm_primal_linear_objective
    elif optimality_norm == L_2:
        primal_err, primal_base = ci.l2_primal_residual, ot_cache.l2_norm_primal_right_hand_side
        dual_err, dual_base = ci.l2_dual_residual, ot_cache.l2_norm_primal_linear_objective
    else:
        raise ValueError("Unknown optimality_norm")
    return (dual_err < abs_tol + rel_tol * dual_base and primal_err < abs_tol + rel_tol * primal_base
            and gap < abs_tol + rel_tol * abs_obj)


def check_termination_criteria(criteria: TerminationCriteria, ot_cache: CachedOTInfo,
                               iteration_stats: IterationStats) -> Optional[TerminationReason]:
    """termination.jl:140-166; returns None where the reference returns false."""
    for ci in iteration_stats.convergence_information:
        if optimality_criteria_met(criteria.optimality_norm, criteria.eps_optimal_absolute,
                                   criteria.eps_optimal_relative, ci, ot_cache):
            return TerminationReason.TERMINATION_REASON_OPTIMAL
    if iteration_stats.iteration_number >= criteria.iteration_limit:
        return TerminationReason.TERMINATION_REASON_ITERATION_LIMIT
    if iteration_stats.cumulative_kkt_matrix_passes >= criteria.kkt_matrix_pass_limit:
        return TerminationReason.TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT
    if iteration_stats.cumulative_time_sec >= criteria.time_sec_limit:
        return TerminationReason.TERMINATION_REASON_TIME_LIMIT
    return None
