"""PDHG driver: restates `optimize` (src/solver_gpu.jl:413-720) over the
C-ABI of libpdot_hip.so.

Evaluation iterations (every termination_evaluation_frequency iterations,
the first ten, iteration_limit+1 and after a numerical error,
solver_gpu.jl:529-532) evaluate the point and its average on the device in
the same streaming pass that speculatively performs the next PDHG step; the
host then checks termination, runs the restart scheme and the primal-weight
update, and either keeps the speculative step or discards it.  Between
evaluations the device runs the steps and their step-size recurrence with no
host round trip.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import List, Optional, Union

import numpy as np

from . import _lib
from .io import OptimalTransportProblem
from .saddle_point import (KKTPieces, RestartParameters, compute_new_primal_weight, create_last_restart_info,
                           run_restart_scheme, select_initial_primal_weight)
from .solve_log import (ConvergenceInformation, IterationStats, PointType, RestartChoice, TerminationReason,
                        termination_reason_to_string)
from .termination import CachedOTInfo, TerminationCriteria, cached_ot_info, check_termination_criteria


@dataclass(frozen=True)
class AdaptiveStepsizeParams:
    """solver_gpu.jl:2-5"""
    reduction_exponent: float
    growth_exponent: float


@dataclass(frozen=True)
class ConstantStepsizeParams:
    """solver_gpu.jl:7"""


@dataclass
class PdhgParameters:
    """solver_gpu.jl:9-21"""
    primal_importance: float
    scale_invariant_initial_primal_weight: bool
    verbosity: int
    record_iteration_stats: bool
    termination_evaluation_frequency: int
    termination_criteria: TerminationCriteria
    restart_params: RestartParameters
    step_size_policy_params: Union[AdaptiveStepsizeParams, ConstantStepsizeParams]


@dataclass
class SaddlePointOutput:
    """saddle_point_gpu.jl:2-34"""
    primal_solution: np.ndarray
    dual_source_solution: np.ndarray
    dual_target_solution: np.ndarray
    termination_reason: TerminationReason
    termination_string: str
    iteration_count: int
    iteration_stats: List[IterationStats] = field(default_factory=list)


def define_norms(step_size: float, primal_weight: float):
    """solver_gpu.jl:51-56"""
    return 1 / step_size * primal_weight, 1 / step_size / primal_weight


def convergence_information(ps, ot_cache: CachedOTInfo, eps_ratio: float,
                            candidate_type: PointType) -> ConvergenceInformation:
    """compute_convergence_information (utils_gpu.jl:68-129) from the device
    reductions of one point."""
    ci = ConvergenceInformation()
    ci.primal_objective = ps.primal_objective
    ci.l_inf_primal_residual = ps.p_inf
    ci.l2_primal_residual = math.sqrt(ps.p_sq)
    ci.relative_l_inf_primal_residual = ci.l_inf_primal_residual / (eps_ratio + ot_cache.l_inf_norm_primal_right_hand_side)
    ci.relative_l2_primal_residual = ci.l2_primal_residual / (eps_ratio + ot_cache.l2_norm_primal_right_hand_side)
    ci.l_inf_primal_variable = ps.x_inf
    ci.l2_primal_variable = math.sqrt(ps.x_sq)
    ci.dual_objective = ps.dual_objective
    ci.l_inf_dual_residual = ps.rc_inf
    ci.l2_dual_residual = math.sqrt(ps.rc_sq)
    ci.relative_l_inf_dual_residual = ci.l_inf_dual_residual / (eps_ratio + ot_cache.l_inf_norm_primal_linear_objective)
    ci.relative_l2_dual_residual = ci.l2_dual_residual / (eps_ratio + ot_cache.l2_norm_primal_linear_objective)
    # [u; u], not [u; v] (utils_gpu.jl:115-116)
    ci.l_inf_dual_variable = ps.u_inf
    ci.l2_dual_variable = math.sqrt(ps.u_sq + ps.u_sq)
    ci.corrected_dual_objective = ps.dual_objective if ps.rc_inf == 0.0 else -math.inf
    gap = abs(ci.primal_objective - ci.dual_objective)
    abs_obj = abs(ci.primal_objective) + abs(ci.dual_objective)
    ci.relative_optimality_gap = gap / (eps_ratio + abs_obj)
    ci.candidate_type = candidate_type
    return ci


def relative_kkt_error(ci: ConvergenceInformation) -> float:
    """max(rel_P, rel_D, rel_gap) in the L2 norm: the quantity the default
    termination test compares against the tolerance."""
    return max(ci.relative_l2_primal_residual, ci.relative_l2_dual_residual, ci.relative_optimality_gap)


def is_evaluation_iteration(iteration: int, frequency: int, iteration_limit: int) -> bool:
    """solver_gpu.jl:529-531 (the numerical-error clause is handled by the loop)."""
    return (iteration - 1) % frequency == 0 or iteration == iteration_limit + 1 or iteration <= 10


def next_evaluation_iteration(iteration: int, frequency: int, iteration_limit: int) -> int:
    if iteration + 1 <= 10:
        return iteration + 1
    nxt = ((iteration - 1) // frequency + 1) * frequency + 1
    if iteration < iteration_limit + 1 < nxt:
        nxt = iteration_limit + 1
    return nxt


def print_to_screen_this_iteration(termination_reason, iteration: int, verbosity: int, frequency: int) -> bool:
    """utils_gpu.jl:184-215"""
    if verbosity >= 2:
        if termination_reason is None:
            num_of_evaluations = (iteration - 1) / frequency
            if verbosity >= 9:
                display_frequency = 1
            elif verbosity >= 6:
                display_frequency = 3
            elif verbosity >= 5:
                display_frequency = 10
            elif verbosity >= 4:
                display_frequency = 20
            elif verbosity >= 3:
                display_frequency = 50
            else:
                return iteration == 1
            return math.fmod(num_of_evaluations, display_frequency) == 0
        return True
    return False


def display_iteration_stats_heading() -> None:
    """utils_gpu.jl:218-243"""
    print(f"{'runtime':<24} | {'residuals':<26} | {' solution information':<26} | {'relative residuals':<23} |")
    print(f"{'#iter':<7} {'#kkt':<8} {'seconds':<7} | {'pr norm':<8} {'du norm':<8}  {'gap':<7} | "
          f"{' pr obj':<9} {'pr norm':<8} {'du norm':<7} | {'rel pr':<7} {'rel du':<7} {'rel gap':<7} |")


def display_iteration_stats(stats: IterationStats) -> None:
    """utils_gpu.jl:249-281"""
    if stats.convergence_information:
        ci = stats.convergence_information[0]
        gap = f"{ci.primal_objective - ci.dual_objective:.1e}".rjust(8)
        pobj = f"{ci.primal_objective:.1e}".rjust(8)
        print(f"{str(stats.iteration_number):<6}  {stats.cumulative_kkt_matrix_passes:.1e}  "
              f"{stats.cumulative_time_sec:.1e} | {ci.l2_primal_residual:.1e}  {ci.l2_dual_residual:.1e}  {gap} | "
              f"{pobj}  {ci.l2_primal_variable:.1e}  {ci.l2_dual_variable:.1e} | "
              f"{ci.relative_l2_primal_residual:.1e} {ci.relative_l2_dual_residual:.1e} "
              f"{ci.relative_optimality_gap:.1e} |")
    else:
        print(f"{str(stats.iteration_number):<6}  {stats.cumulative_kkt_matrix_passes:.1e}  "
              f"{stats.cumulative_time_sec:.1e}")


@dataclass
class SolveProfile:
    """Timing side-channel of one optimize call (not part of the reference
    output): wall time from after the upload to the terminating evaluation
    (where the reference reads cumulative_time_sec, solver_gpu.jl:501-620),
    the device-to-host copy of the output after it, and the number of
    evaluations."""
    upload_sec: float = 0.0
    solve_sec: float = 0.0
    output_sec: float = 0.0
    evaluations: int = 0
    restarts: int = 0
    take_step_calls: int = 0


def initial_step_and_weight(params: PdhgParameters, nrow: int, ncol: int, ot_cache: CachedOTInfo):
    """Initial step size, primal weight and step policy (solver_gpu.jl:476-496)."""
    sp = params.step_size_policy_params
    if isinstance(sp, AdaptiveStepsizeParams):
        policy, re, ge = _lib.POLICY_ADAPTIVE, float(sp.reduction_exponent), float(sp.growth_exponent)
        step_size = 1.0
    else:
        policy, re, ge = _lib.POLICY_CONSTANT, 0.0, 0.0
        desired_relative_error = 0.2
        step_size = (1 - desired_relative_error) / math.sqrt(nrow + ncol)
    if params.scale_invariant_initial_primal_weight:
        primal_weight = select_initial_primal_weight(ot_cache.l2_norm_primal_linear_objective,
                                                     ot_cache.l2_norm_primal_right_hand_side, 1.0, 1.0,
                                                     params.primal_importance, params.verbosity)
    else:
        primal_weight = params.primal_importance
    return policy, re, ge, step_size, primal_weight


class InstanceController:
    """The host half of one solve: at each evaluation, the statistics,
    termination check and restart decision of optimize (solver_gpu.jl:529-696)
    from the device reductions of the current point and its average."""

    KKT_PASSES_PER_TERMINATION_EVALUATION = 2.0

    def __init__(self, params: PdhgParameters, ot_cache: CachedOTInfo, restart_trace: Optional[list] = None):
        self.params = params
        self.restart_trace = restart_trace
        self.ot_cache = ot_cache
        crit = params.termination_criteria
        self.eps_ratio = crit.eps_optimal_absolute / crit.eps_optimal_relative
        self.eval_passes = 0.0
        self.last_restart_info = None
        self.iteration_stats: List[IterationStats] = []
        self.evaluations = 0
        self.restarts = 0

    def evaluate(self, cur, avg, st, iteration: int, elapsed: float, time_basic: float):
        """Returns (termination_reason or None, stats)."""
        self.eval_passes += self.KKT_PASSES_PER_TERMINATION_EVALUATION
        self.evaluations += 1
        if self.last_restart_info is None:
            self.last_restart_info = create_last_restart_info(KKTPieces.from_point_stats(cur))
        numerical_error = bool(st.numerical_error)
        point = cur if (numerical_error or st.avg_count == 0) else avg
        stats = IterationStats()
        stats.iteration_number = iteration
        stats.cumulative_kkt_matrix_passes = st.cumulative_kkt_passes + self.eval_passes
        stats.cumulative_time_sec = elapsed
        stats.convergence_information = [convergence_information(point, self.ot_cache, self.eps_ratio,
                                                                 PointType.POINT_TYPE_AVERAGE_ITERATE)]
        stats.step_size = st.step_size
        stats.primal_weight = st.primal_weight
        stats.method_specific_stats = {"time_spent_doing_basic_algorithm": time_basic}
        reason = check_termination_criteria(self.params.termination_criteria, self.ot_cache, stats)
        if numerical_error and reason is None:
            reason = TerminationReason.TERMINATION_REASON_NUMERICAL_ERROR
        if self.params.record_iteration_stats or reason is not None:
            self.iteration_stats.append(stats)
        return reason, stats

    def restart(self, cur, avg, st, iteration: int, stats: IterationStats):
        """run_restart_scheme + compute_new_primal_weight; returns
        (decision, new primal weight)."""
        primal_norm_params, dual_norm_params = define_norms(st.step_size, st.primal_weight)
        rp = self.params.restart_params
        last_point = self.last_restart_info.point
        decision = run_restart_scheme(int(st.avg_count), KKTPieces.from_point_stats(cur),
                                      KKTPieces.from_point_stats(avg) if st.avg_count > 0 else None,
                                      self.last_restart_info, iteration - 1, primal_norm_params, dual_norm_params,
                                      st.primal_weight, rp)
        stats.restart_used = decision.choice
        weight = st.primal_weight
        if decision.choice != RestartChoice.RESTART_CHOICE_NO_RESTART:
            self.restarts += 1
            weight = compute_new_primal_weight(self.last_restart_info, st.primal_weight,
                                               rp.primal_weight_update_smoothing, self.params.verbosity)
        if self.restart_trace is not None:
            self.restart_trace.append(restart_record(decision, cur, avg, last_point, st, iteration, weight))
        return decision, weight


def _pieces_dict(pieces: Optional[KKTPieces]) -> dict:
    keys = ("l2_primal_residual", "l2_dual_residual", "primal_objective", "dual_objective", "x_norm", "y_norm")
    if pieces is None:
        return {k: math.nan for k in keys}
    return {k: getattr(pieces, k) for k in keys}


def restart_record(decision, cur, avg, last_point: KKTPieces, st, iteration: int, weight_after: float) -> dict:
    """One evaluation's restart decision with everything it read, for logs and
    tests (the fields of the oracle's restart trace)."""
    return {"iteration": int(iteration), "restart_used": int(decision.choice), "artificial": int(decision.artificial),
            "adaptive_checked": int(decision.adaptive_checked), "reset_to_average": int(decision.reset_to_average),
            "avg_count": int(st.avg_count), "iterations_completed": int(iteration) - 1,
            "step_size": st.step_size, "primal_weight_before": st.primal_weight, "primal_weight_after": weight_after,
            "kkt_current": decision.kkt_current, "kkt_average": decision.kkt_average,
            "kkt_last": decision.kkt_last, "ratio": decision.ratio,
            "last_trial_before": decision.last_trial_before, "last_trial_after": decision.last_trial_after,
            "last_x_norm": decision.last_x_norm, "last_y_norm": decision.last_y_norm,
            "primal_distance": decision.primal_distance, "dual_distance": decision.dual_distance,
            "current": _pieces_dict(KKTPieces.from_point_stats(cur)),
            "average": _pieces_dict(KKTPieces.from_point_stats(avg) if st.avg_count > 0 else None),
            "last": _pieces_dict(last_point)}


def optimize(params: PdhgParameters, original_problem: OptimalTransportProblem, *,
             storage: int = _lib.STORAGE_F64, device: int = 0, warm_start=None,
             solver: Optional[_lib.DeviceSolver] = None, profile: Optional[SolveProfile] = None,
             use_graphs: bool = False, restart_trace: Optional[list] = None) -> SaddlePointOutput:
    """PDOT.optimize (solver_gpu.jl:413-720) on an MI355X.

    storage: _lib.STORAGE_F64 (the reference's precision) or
    _lib.STORAGE_F32 (C, X and the average stored in FP32; duals, sums and
    scalars in FP64).  warm_start: optional (u0, v0) initial duals; X starts at
    0 and G at (C - u0) - v0.  solver: reuse a DeviceSolver whose problem is
    already uploaded (the upload is then skipped).  restart_trace: a list
    that receives one record per non-terminating evaluation (restart_record)."""
    from .io import PointCloudProblem
    points = isinstance(original_problem, PointCloudProblem)
    nrow, ncol = original_problem.num_row, original_problem.num_col
    if params.primal_importance <= 0 or not math.isfinite(params.primal_importance):
        raise ValueError("primal_importance must be positive and finite")
    if params.termination_evaluation_frequency <= 0:
        raise ValueError("termination_evaluation_frequency must be positive")

    t_up = time.time()
    if solver is None:
        solver = _lib.DeviceSolver(nrow, ncol, storage=storage, device=device)
        if points:
            solver.upload_points(original_problem)
        else:
            solver.upload(original_problem.cost_matrix, original_problem.source_distribution,
                          original_problem.target_distribution)
    elif (solver.n, solver.m) != (nrow, ncol):
        raise ValueError("solver was created for another problem size")
    solver.use_graphs(use_graphs)
    upload_sec = time.time() - t_up
    # termination.jl:94-101, from the device upload pass (no host pass over C)
    ot_cache = solver.cached_ot_info()

    policy, re, ge, step_size, primal_weight = initial_step_and_weight(params, nrow, ncol, ot_cache)
    u0 = v0 = None
    if warm_start is not None:
        u0, v0 = warm_start
    solver.set_start(u0, v0, policy, re, ge, step_size, primal_weight)

    ctl = InstanceController(params, ot_cache, restart_trace)
    start_time = time.time()
    time_basic = 0.0
    iteration_limit = params.termination_criteria.iteration_limit
    freq = params.termination_evaluation_frequency
    if params.verbosity >= 2:
        display_iteration_stats_heading()

    iteration = 1
    while True:
        # `iteration` is an evaluation iteration: evaluate the state before its step
        t0 = time.time()
        cur, avg, st = solver.eval()
        time_basic += time.time() - t0
        reason, stats = ctl.evaluate(cur, avg, st, iteration, time.time() - start_time, time_basic)
        if print_to_screen_this_iteration(reason, iteration, params.verbosity, freq):
            display_iteration_stats(stats)
        if reason is not None:
            solve_sec = time.time() - start_time
            solver.discard()
            which = _lib.POINT_CURRENT if st.avg_count == 0 else _lib.POINT_AVERAGE
            solver.output_point = which  # the returned point stays resident (rounding on the device)
            X, u, v = solver.download(which)
            if profile is not None:
                profile.upload_sec = upload_sec
                profile.solve_sec = solve_sec
                profile.output_sec = time.time() - start_time - solve_sec
                profile.evaluations = ctl.evaluations
                profile.restarts = ctl.restarts
                profile.take_step_calls = int(st.take_step_calls)
            if params.verbosity >= 1:
                print(f"Terminated after {iteration} iterations: {termination_reason_to_string(reason)}")
            return SaddlePointOutput(X, u, v, reason, termination_reason_to_string(reason), iteration - 1,
                                     ctl.iteration_stats)
        decision, weight = ctl.restart(cur, avg, st, iteration, stats)
        t0 = time.time()
        nxt = next_evaluation_iteration(iteration, freq, iteration_limit)
        if decision.choice != RestartChoice.RESTART_CHOICE_NO_RESTART:
            solver.restart(decision.to_average, weight)
            count = nxt - iteration  # the speculative step was discarded: redo it
        else:
            count = nxt - iteration - 1  # the speculative step is committed first
        st = solver.run_steps(count)
        time_basic += time.time() - t0
        if st.numerical_error:
            iteration = int(st.error_call) + 1
        else:
            iteration = nxt
