"""CPU tests of the output formats: Julia's Float64 text (the library's
host formatter and the Python one used for the JSON logs), the vector
files of write_vector_to_file (scripts/solve.jl:5-11) and the SolveLog JSON
of solve_instance_and_output (scripts/solve.jl:88-110).  No GPU call."""
import ctypes
import json
import math

import numpy as np
import pytest

from pdot_amd import _lib
from pdot_amd.output import julia_float_text, solve_log_json, write_vector_to_file
from pdot_amd.solve_log import (ConvergenceInformation, IterationStats, PointType, RestartChoice, SolveLog,
                                TerminationReason)

# How Julia prints these Float64 values (Base.Ryu.writeshortest through
# print/println): positional for 1e-4 <= |x| < 1e6, else d.ddde<E> with at
# least one fraction digit and no exponent padding or plus sign.
JULIA_TEXT = [(0.0, "0.0"), (-0.0, "-0.0")] + list({
    1.0: "1.0", -1.0: "-1.0", 0.1: "0.1", 0.5: "0.5",
    1e-5: "1.0e-5", 1.5e-5: "1.5e-5", 0.0001: "0.0001", 0.00012: "0.00012",
    100000.0: "100000.0", 123456.0: "123456.0", 999999.9: "999999.9",
    1e6: "1.0e6", 1234567.0: "1.234567e6", 1e16: "1.0e16",
    1 / 3: "0.3333333333333333", math.pi: "3.141592653589793", 0.1 + 0.2: "0.30000000000000004",
    2.0 ** -52: "2.220446049250313e-16", 5e-324: "5.0e-324",
    1.7976931348623157e308: "1.7976931348623157e308", -2.5e-300: "-2.5e-300",
    math.inf: "Inf", -math.inf: "-Inf",
}.items())


def _lib_text(x):
    lib = _lib.load_library()
    fn = lib.pdot_format_julia_float
    fn.restype, fn.argtypes = _lib.SIGNATURES["pdot_format_julia_float"]
    buf = ctypes.create_string_buffer(40)
    k = fn(float(x), buf, 40)
    assert k > 0
    return buf.value.decode()


@pytest.mark.parametrize("x,text", JULIA_TEXT)
def test_julia_float_known_answers(x, text):
    assert math.copysign(1, x) == (-1 if text.startswith("-") else 1)
    assert julia_float_text(x) == text
    assert _lib_text(x) == text


def test_nan_text():
    assert julia_float_text(math.nan) == "NaN" and _lib_text(math.nan) == "NaN"


def test_formatters_agree_and_round_trip(tmp_path):
    """Random bit patterns over the whole exponent range plus decimal-looking
    values: the C++ and Python formatters agree, and every text reads back to
    the same double."""
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2**63 - 1, 20000, dtype=np.int64).view(np.float64)
    bits = bits[np.isfinite(bits)]
    dec = np.array([round(float(v) * 10.0 ** int(e), int(d))
                    for v, e, d in zip(rng.random(5000), rng.integers(-8, 9, 5000), rng.integers(0, 8, 5000))])
    x = np.concatenate([bits, -bits[:100], dec, [0.0, -0.0, 1e-4, 9.999999999999999e-5, 999999.9999999999, 1e6]])
    path = tmp_path / "v.txt"
    write_vector_to_file(str(path), x)
    lines = open(path).read().splitlines()
    assert len(lines) == x.size
    for xi, line in zip(x, lines):
        assert line == julia_float_text(xi)
        assert float(line) == xi and math.copysign(1, float(line)) == math.copysign(1, xi)


def test_write_vector_matrix_is_column_major(tmp_path):
    X = np.arange(12, dtype=np.float64).reshape(3, 4) / 8
    path = tmp_path / "m.txt"
    write_vector_to_file(str(path), X)
    assert open(path).read().splitlines() == [julia_float_text(v) for v in X.ravel(order="F")]
    write_vector_to_file(str(path), np.zeros(0))
    assert open(path).read() == ""


def test_write_vector_reports_errors(tmp_path):
    with pytest.raises(OSError, match="cannot open"):
        write_vector_to_file(str(tmp_path / "no" / "such" / "dir.txt"), np.ones(3))


def test_solve_log_json_layout():
    ci = ConvergenceInformation(candidate_type=PointType.POINT_TYPE_AVERAGE_ITERATE, primal_objective=1e-5,
                                dual_objective=math.inf, l2_dual_variable=123456.0)
    st = IterationStats(iteration_number=129, convergence_information=[ci], cumulative_kkt_matrix_passes=260.0,
                        restart_used=RestartChoice.RESTART_CHOICE_NO_RESTART, step_size=0.5, primal_weight=2.0)
    log = SolveLog(instance_name="x", command_line_invocation="solve.jl x",
                   termination_reason=TerminationReason.TERMINATION_REASON_OPTIMAL, termination_string="OPTIMAL",
                   iteration_count=128, solve_time_sec=0.25, solution_stats=st,
                   solution_type=PointType.POINT_TYPE_AVERAGE_ITERATE)
    text = solve_log_json(log)
    assert '"primal_objective":1.0e-5' in text
    assert '"dual_objective":Infinity' in text
    assert '"l2_dual_variable":123456.0' in text
    assert '"cumulative_kkt_matrix_passes":260.0' in text
    assert '"iteration_number":129' in text
    d = json.loads(text)
    assert list(d) == ["instance_name", "command_line_invocation", "termination_reason", "termination_string",
                       "iteration_count", "solve_time_sec", "solution_stats", "solution_type", "iteration_stats"]
    assert d["termination_reason"] == "TERMINATION_REASON_OPTIMAL"
    assert d["solution_stats"]["restart_used"] == "RESTART_CHOICE_NO_RESTART"
    assert d["solution_stats"]["convergence_information"][0]["candidate_type"] == "POINT_TYPE_AVERAGE_ITERATE"
    assert d["solution_stats"]["convergence_information"][0]["dual_objective"] == math.inf
