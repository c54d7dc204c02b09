"""pdot_amd: MI355X-native PDOT (restarted adaptive PDHG for discrete optimal
transport).  The API mirrors the PDOT.jl module (src/PDOT.jl): problem type,
parameter structs and constructors, `optimize`, the solve logs and the
script-level `solve_instance_and_output`.  All n*m work runs in
libpdot_hip.so (hand-written gfx950 HIP kernels behind a C ABI,
include/pdot.h); this package is the host control logic."""
from __future__ import annotations

from ._lib import (STORAGE_F32, STORAGE_F64, DeviceSolver, PdotError, load_library)
from .io import OptimalTransportProblem, PointCloudProblem
from .saddle_point import (ADAPTIVE_KKT, FIXED_FREQUENCY, KKT_GREEDY, NO_RESTART_TO_CURRENT, NO_RESTARTS,
                           RestartParameters, RestartScheme, RestartToCurrentMetric, compute_new_primal_weight,
                           construct_restart_parameters, select_initial_primal_weight)
from .solve_log import (ConvergenceInformation, IterationStats, PointType, RestartChoice, SolveLog,
                        TerminationReason, termination_reason_to_string, to_json)
from .solve_log import *  # noqa: F401,F403  (enum instances, as the Julia module exposes them)
from .solver import (AdaptiveStepsizeParams, ConstantStepsizeParams, PdhgParameters, SaddlePointOutput,
                     SolveProfile, optimize, relative_kkt_error)
from .termination import (L_2, L_INF, CachedOTInfo, OptimalityNorm, TerminationCriteria, cached_ot_info,
                          check_termination_criteria, construct_termination_criteria, optimality_criteria_met,
                          validate_termination_criteria)


def default_parameters(tolerance: float = 1e-4, *, iteration_limit: int = 2**31 - 1,
                       time_sec_limit: float = float("inf"), verbosity: int = 0,
                       record_iteration_stats: bool = True) -> PdhgParameters:
    """The parameters scripts/solve.jl:58-86 builds: ADAPTIVE_KKT restarts
    with KKT_GREEDY (1000, 0.36, 0.1, 0.9, 0.5), L2 optimality at
    `tolerance` absolute and relative, evaluation every 128 iterations,
    AdaptiveStepsizeParams(2, 1), primal_importance 1."""
    restart_params = construct_restart_parameters(ADAPTIVE_KKT, KKT_GREEDY, 1000, 0.36, 0.1, 0.9, 0.5)
    termination_params = construct_termination_criteria(
        optimality_norm=L_2, eps_optimal_absolute=tolerance, eps_optimal_relative=tolerance,
        time_sec_limit=time_sec_limit, iteration_limit=iteration_limit, kkt_matrix_pass_limit=float("inf"))
    return PdhgParameters(1.0, False, verbosity, record_iteration_stats, 128, termination_params, restart_params,
                          AdaptiveStepsizeParams(2, 1))


from .output import rounding, solve_instance_and_output, write_vector_to_file  # noqa: E402
