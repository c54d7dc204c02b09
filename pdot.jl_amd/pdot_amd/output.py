"""Solution output: feasibility rounding and the files of
solve_instance_and_output (scripts/solve.jl:48-130).

    rounding(problem, primal_solution)         src/utils_gpu.jl:296-318
    write_vector_to_file(filename, vector)     scripts/solve.jl:5-11
    solve_instance_and_output(output_directory, instance_name, instance; tolerance)

The rounding runs on the GPU (libpdot_hip.so: pdot_round for a plan still
resident in a solver, pdot_round_matrix for a host matrix); the text writer is
the library's threaded host formatter.  Numbers are written as Julia prints a
Float64 (shortest round-trip digits, "1.0e-5", "0.0001", "123456.0").

Where the reference script cannot run as written, the intent is kept and the
choice recorded here:
  * it reads the globals output_dir and time_sec_limit: the directory argument
    and a time_sec_limit keyword (default Inf) are used;
  * warm_up reads the globals tolerance and warmup_termination_params (the
    local is named termination_params_warmup): the warm-up solve uses the
    instance's tolerance and its 100-iteration limit;
  * it writes output.dual_solution, a field SaddlePointOutput does not have
    (saddle_point_gpu.jl:2-34): the dual file holds [u; v], the duals of the
    source then the target constraints.
"""
from __future__ import annotations

import contextlib
import ctypes
import gzip
import io
import json
import math
import os
import sys
from typing import Optional

import numpy as np

from . import _lib
from .io import OptimalTransportProblem
from .solve_log import PointType, SolveLog, _jsonable
from .solver import PdhgParameters, optimize

_ERRLEN = 512


def julia_float_text(x: float) -> str:
    """Julia's print of a Float64: shortest round-trip digits, positional for
    1e-4 <= |x| < 1e6 and "d.ddde<E>" otherwise (Base.Ryu.writeshortest)."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    x = abs(x)
    if x == 0.0:
        return sign + "0.0"
    r = repr(x)  # shortest round-trip digits
    mant, _, exp = r.partition("e")
    e10 = int(exp) if exp else 0
    ip, _, fp = mant.partition(".")
    digits = (ip + fp).lstrip("0")
    lead = len(ip + fp) - len((ip + fp).lstrip("0"))
    E = len(ip) - 1 - lead + e10  # value = d1.d2... x 10^E
    digits = digits.rstrip("0") or "0"
    k, pt = len(digits), E + 1
    if -4 < pt <= 6:
        if pt <= 0:
            return sign + "0." + "0" * (-pt) + digits
        if pt < k:
            return sign + digits[:pt] + "." + digits[pt:]
        return sign + digits + "0" * (pt - k) + ".0"
    return sign + digits[0] + "." + (digits[1:] or "0") + "e" + str(E)


def _json_text(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if v is None:
        return "null"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        if math.isnan(v):
            return "NaN"
        if math.isinf(v):
            return "Infinity" if v > 0 else "-Infinity"
        return julia_float_text(v)
    if isinstance(v, str):
        return json.dumps(v)
    if isinstance(v, dict):
        return "{" + ",".join(_json_text(str(k)) + ":" + _json_text(x) for k, x in v.items()) + "}"
    if isinstance(v, (list, tuple)):
        return "[" + ",".join(_json_text(x) for x in v) + "]"
    raise TypeError(f"cannot write {type(v)} as JSON")


def solve_log_json(log: SolveLog) -> str:
    """JSON3.write(log, allow_inf=true): fields in declaration order, enums by
    name, finite numbers in Julia's float text, infinities as Infinity."""
    return _json_text(_jsonable(log))


def write_vector_to_file(filename: str, vector) -> None:
    """scripts/solve.jl:5-11: one element per line; a matrix in column-major
    order."""
    lib = _lib.load_library()
    x = np.asarray(vector, dtype=np.float64)
    x = x.ravel(order="F") if x.ndim > 1 else np.ascontiguousarray(x)
    x = np.ascontiguousarray(x)
    err = ctypes.create_string_buffer(_ERRLEN)
    rc = lib.pdot_write_julia_vector(os.fsencode(filename), x.ctypes.data_as(_lib._c_double_p), x.size, err, _ERRLEN)
    if rc != _lib.PDOT_OK:
        raise OSError(err.value.decode() or f"pdot_write_julia_vector failed ({rc})")


def rounding(problem: OptimalTransportProblem, primal_solution: np.ndarray, device: int = 0) -> np.ndarray:
    """src/utils_gpu.jl:296-318 on the GPU.  Like the reference, the plan is
    overwritten with its rounding and returned; an array that is not a
    column-major float64 n x m matrix is converted first, and the converted
    copy is the one rounded and returned."""
    C = problem.cost_matrix
    n, m = C.shape
    X = primal_solution
    if not (isinstance(X, np.ndarray) and X.dtype == np.float64 and X.flags.f_contiguous and X.flags.writeable):
        X = np.asfortranarray(X, dtype=np.float64).copy(order="F")
    if X.shape != (n, m):
        raise ValueError(f"primal_solution shape {X.shape} != {(n, m)}")
    a = np.ascontiguousarray(problem.source_distribution, dtype=np.float64)
    b = np.ascontiguousarray(problem.target_distribution, dtype=np.float64)
    lib = _lib.load_library()
    err = ctypes.create_string_buffer(_ERRLEN)
    rc = lib.pdot_round_matrix(int(device), n, m, a.ctypes.data_as(_lib._c_double_p), b.ctypes.data_as(_lib._c_double_p),
                               X.ctypes.data_as(_lib._c_double_p), err, _ERRLEN)
    if rc != _lib.PDOT_OK:
        raise _lib.PdotError(f"pdot_round_matrix failed ({rc}): {err.value.decode()}")
    return X


def solve_parameters(tolerance: float, *, verbosity: int = 2, time_sec_limit: float = math.inf,
                     iteration_limit: int = 2**31 - 1) -> PdhgParameters:
    """The parameters of scripts/solve.jl:58-86 (verbosity 2, iteration
    statistics recorded)."""
    from . import default_parameters
    return default_parameters(tolerance, iteration_limit=iteration_limit, time_sec_limit=time_sec_limit,
                              verbosity=verbosity, record_iteration_stats=True)


def warm_up(instance: OptimalTransportProblem, tolerance: float = 1e-4, *, storage: int = _lib.STORAGE_F64,
            device: int = 0) -> None:
    """scripts/solve.jl:13-46: a 100-iteration solve, output discarded."""
    optimize(solve_parameters(tolerance, verbosity=0, iteration_limit=100), instance, storage=storage, device=device)


def solve_instance_and_output(output_directory: str, instance_name: str, instance: OptimalTransportProblem, *,
                              tolerance: float = 1.0e-4, time_sec_limit: float = math.inf,
                              storage: int = _lib.STORAGE_F64, device: int = 0, do_warm_up: bool = True,
                              verbosity: int = 2, command_line_invocation: Optional[str] = None):
    """scripts/solve.jl:48-130: solve to `tolerance` and write
    <name>_summary.json, <name>_full_log.json.gz, <name>_primal.txt,
    <name>_primal_rounded.txt and <name>_dual.txt under output_directory.
    The rounded plan is computed on the device from the resident solution.
    Returns the SaddlePointOutput."""
    os.makedirs(output_directory, exist_ok=True)
    if do_warm_up:
        with contextlib.redirect_stdout(io.StringIO()):
            warm_up(instance, tolerance, storage=storage, device=device)

    n, m = instance.cost_matrix.shape
    solver = _lib.DeviceSolver(n, m, storage=storage, device=device)
    try:
        solver.upload(instance.cost_matrix, instance.source_distribution, instance.target_distribution)
        output = optimize(solve_parameters(tolerance, verbosity=verbosity, time_sec_limit=time_sec_limit), instance,
                          storage=storage, device=device, solver=solver)

        log = SolveLog()
        log.instance_name = instance_name
        log.command_line_invocation = (" ".join(sys.argv) if command_line_invocation is None
                                       else command_line_invocation)
        log.termination_reason = output.termination_reason
        log.termination_string = output.termination_string
        log.iteration_count = output.iteration_count
        log.solve_time_sec = output.iteration_stats[-1].cumulative_time_sec
        log.solution_stats = output.iteration_stats[-1]
        log.solution_type = PointType.POINT_TYPE_AVERAGE_ITERATE

        base = os.path.join(output_directory, instance_name)
        with open(base + "_summary.json", "w") as f:
            f.write(solve_log_json(log))
        log.iteration_stats = output.iteration_stats
        with gzip.open(base + "_full_log.json.gz", "wt") as f:
            f.write(solve_log_json(log))

        write_vector_to_file(base + "_primal.txt", output.primal_solution)
        write_vector_to_file(base + "_primal_rounded.txt", solver.round(solver.output_point))
        write_vector_to_file(base + "_dual.txt",
                             np.concatenate([output.dual_source_solution, output.dual_target_solution]))
    finally:
        solver.close()
    return output
