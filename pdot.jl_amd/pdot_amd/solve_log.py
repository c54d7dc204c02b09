"""Solver records: enums, ConvergenceInformation, IterationStats, SolveLog.

Restates src/solve_log.jl (field names, order, defaults and enum values) and
its JSON form (StructTypes.Mutable objects written by JSON3 with enums as
their names and allow_inf, src/solve_log.jl:330-332, scripts/solve.jl:103).
"""
from __future__ import annotations

import enum
import json
import math
from dataclasses import dataclass, field
from typing import Dict, List


class RestartChoice(enum.IntEnum):
    """solve_log.jl:15-20"""
    RESTART_CHOICE_UNSPECIFIED = 0
    RESTART_CHOICE_NO_RESTART = 1
    RESTART_CHOICE_WEIGHTED_AVERAGE_RESET = 2
    RESTART_CHOICE_RESTART_TO_AVERAGE = 3


class PointType(enum.IntEnum):
    """solve_log.jl:35-41"""
    POINT_TYPE_UNSPECIFIED = 0
    POINT_TYPE_CURRENT_ITERATE = 1
    POINT_TYPE_ITERATE_DIFFERENCE = 2
    POINT_TYPE_AVERAGE_ITERATE = 3
    POINT_TYPE_NONE = 4


class TerminationReason(enum.IntEnum):
    """solve_log.jl:245-254"""
    TERMINATION_REASON_UNSPECIFIED = 0
    TERMINATION_REASON_OPTIMAL = 1
    TERMINATION_REASON_TIME_LIMIT = 2
    TERMINATION_REASON_ITERATION_LIMIT = 3
    TERMINATION_REASON_KKT_MATRIX_PASS_LIMIT = 4
    TERMINATION_REASON_NUMERICAL_ERROR = 5
    TERMINATION_REASON_INVALID_PROBLEM = 6
    TERMINATION_REASON_OTHER = 7


# module-level aliases, as Julia exports enum instances into the module
globals().update({e.name: e for cls in (RestartChoice, PointType, TerminationReason) for e in cls})

CONVERGENCE_FIELDS = (
    "candidate_type", "primal_objective", "dual_objective", "corrected_dual_objective",
    "l_inf_primal_residual", "l2_primal_residual", "l_inf_dual_residual", "l2_dual_residual",
    "relative_l_inf_primal_residual", "relative_l2_primal_residual", "relative_l_inf_dual_residual",
    "relative_l2_dual_residual", "relative_optimality_gap", "l_inf_primal_variable", "l2_primal_variable",
    "l_inf_dual_variable", "l2_dual_variable",
)


@dataclass
class ConvergenceInformation:
    """solve_log.jl:47-151; the zero-argument form fills zeros (:149-151)."""
    candidate_type: PointType = PointType.POINT_TYPE_UNSPECIFIED
    primal_objective: float = 0.0
    dual_objective: float = 0.0
    corrected_dual_objective: float = 0.0
    l_inf_primal_residual: float = 0.0
    l2_primal_residual: float = 0.0
    l_inf_dual_residual: float = 0.0
    l2_dual_residual: float = 0.0
    relative_l_inf_primal_residual: float = 0.0
    relative_l2_primal_residual: float = 0.0
    relative_l_inf_dual_residual: float = 0.0
    relative_l2_dual_residual: float = 0.0
    relative_optimality_gap: float = 0.0
    l_inf_primal_variable: float = 0.0
    l2_primal_variable: float = 0.0
    l_inf_dual_variable: float = 0.0
    l2_dual_variable: float = 0.0


@dataclass
class IterationStats:
    """solve_log.jl:154-228"""
    iteration_number: int = 0
    convergence_information: List[ConvergenceInformation] = field(default_factory=list)
    cumulative_kkt_matrix_passes: float = 0.0
    cumulative_rejected_steps: int = 0
    cumulative_time_sec: float = 0.0
    restart_used: RestartChoice = RestartChoice.RESTART_CHOICE_UNSPECIFIED
    step_size: float = 0.0
    primal_weight: float = 0.0
    method_specific_stats: Dict[str, float] = field(default_factory=dict)


@dataclass
class SolveLog:
    """solve_log.jl:256-327"""
    instance_name: str = ""
    command_line_invocation: str = ""
    termination_reason: TerminationReason = TerminationReason.TERMINATION_REASON_UNSPECIFIED
    termination_string: str = ""
    iteration_count: int = 0
    solve_time_sec: float = 0.0
    solution_stats: IterationStats = field(default_factory=IterationStats)
    solution_type: PointType = PointType.POINT_TYPE_UNSPECIFIED
    iteration_stats: List[IterationStats] = field(default_factory=list)


def _jsonable(x):
    if isinstance(x, enum.IntEnum):
        return x.name
    if isinstance(x, (ConvergenceInformation, IterationStats, SolveLog)):
        return {k: _jsonable(getattr(x, k)) for k in x.__dataclass_fields__}
    if isinstance(x, list):
        return [_jsonable(v) for v in x]
    if isinstance(x, dict):
        return {str(k): _jsonable(v) for k, v in x.items()}
    if isinstance(x, float):
        return x
    return x


def to_json(obj) -> str:
    """JSON text of a record; infinities are written as Infinity/-Infinity
    like JSON3.write(..., allow_inf=true)."""
    return json.dumps(_jsonable(obj), allow_nan=True, separators=(",", ":"))


def termination_reason_to_string(reason: TerminationReason) -> str:
    """termination.jl:168-171: the enum name without TERMINATION_REASON_."""
    return TerminationReason(reason).name[len("TERMINATION_REASON_"):]


def isfinite_or_neg_inf(x: float) -> bool:
    return math.isfinite(x) or x == -math.inf
