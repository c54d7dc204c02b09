"""GPU tests of the output path: feasibility rounding on the device
(pdot_round, pdot_round_matrix) against the oracle's restatement of
src/utils_gpu.jl:296-318, and the files of solve_instance_and_output
(scripts/solve.jl:48-130)."""
import gzip
import json
import os

import numpy as np
import pytest

import oracle as O
import pdot_amd as P
from pdot_amd import _lib, instances
from pdot_amd.output import julia_float_text, solve_instance_and_output

pytestmark = pytest.mark.gpu


def _plan(rng, n, m, zero_row=True):
    X = np.asfortranarray(rng.random((n, m)) * rng.choice([0.5, 2.0], size=(n, 1)) / (n * m))
    if zero_row and n > 2 and m > 2:
        X[n // 2, :] = 0.0
        X[:, m // 2] = 0.0
    a = rng.random(n); a /= a.sum()
    b = rng.random(m); b /= b.sum()
    return a, b, X


@pytest.mark.parametrize("n,m", [(1, 1), (1, 300), (300, 1), (63, 65), (257, 1000), (1500, 777)])
def test_round_matrix_matches_oracle(n, m):
    """Only the summation order differs from the oracle: agreement to
    rounding (1e-13 relative to the largest entry), exact feasibility up to
    rounding, and the input array is overwritten as the reference's is."""
    rng = np.random.default_rng(n * 1000 + m)
    a, b, X = _plan(rng, n, m)
    ref = O.rounding(a, b, X)
    prob = P.OptimalTransportProblem(np.zeros((n, m), order="F"), a, b)
    Xd = X.copy(order="F")
    out = P.rounding(prob, Xd)
    assert out is Xd
    scale = max(np.abs(ref).max(), 1e-300)
    np.testing.assert_allclose(out, ref, rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(out.sum(axis=1), a, atol=1e-14)
    np.testing.assert_allclose(out.sum(axis=0), b, atol=1e-14)


def test_round_matrix_feasible_plan_is_bit_identical():
    """Dyadic feasible plan: every sum is exact in any order, so the GPU
    result equals the input bit for bit (s = t = 1, zero residuals)."""
    rng = np.random.default_rng(9)
    a = rng.integers(1, 64, 700).astype(float)
    b = rng.integers(1, 64, 333).astype(float)
    a /= 2.0 ** np.ceil(np.log2(a.sum())); a[0] += 1.0 - a.sum()
    b /= 2.0 ** np.ceil(np.log2(b.sum())); b[0] += 1.0 - b.sum()
    X = np.asfortranarray(np.outer(a, b))
    prob = P.OptimalTransportProblem(np.zeros(X.shape, order="F"), a, b)
    assert np.array_equal(P.rounding(prob, X.copy(order="F")), X)


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_round_resident_point_matches_oracle(storage):
    """pdot_round on the solver's resident average equals the oracle's
    rounding of the downloaded (widened) point to rounding."""
    prob = instances.uniform_sqeuclid(900, 700, seed=12)
    dev = _lib.DeviceSolver(900, 700, storage=storage)
    dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    out = P.optimize(P.default_parameters(1e-3, iteration_limit=300), prob, storage=storage, solver=dev)
    Xr = dev.round(dev.output_point)
    ref = O.rounding(prob.source_distribution, prob.target_distribution, out.primal_solution)
    np.testing.assert_allclose(Xr, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    X2, _, _ = dev.download(dev.output_point)
    assert np.array_equal(X2, out.primal_solution)  # the resident point is left unchanged
    dev.close()


def test_solve_instance_and_output_files(tmp_path):
    prob = instances.uniform_sqeuclid(120, 90, seed=21)
    out = solve_instance_and_output(str(tmp_path), "inst", prob, tolerance=1e-4, verbosity=0,
                                    command_line_invocation="solve.jl inst")
    names = sorted(os.listdir(tmp_path))
    assert names == ["inst_dual.txt", "inst_full_log.json.gz", "inst_primal.txt", "inst_primal_rounded.txt",
                     "inst_summary.json"]
    assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL

    lines = open(tmp_path / "inst_primal.txt").read().splitlines()
    X = out.primal_solution
    assert len(lines) == X.size
    assert lines[:50] == [julia_float_text(x) for x in X.ravel(order="F")[:50]]
    assert np.array_equal(np.array([float(s) for s in lines]), X.ravel(order="F"))

    rounded = np.array([float(s) for s in open(tmp_path / "inst_primal_rounded.txt").read().splitlines()])
    ref = O.rounding(prob.source_distribution, prob.target_distribution, X).ravel(order="F")
    np.testing.assert_allclose(rounded, ref, rtol=0, atol=1e-13 * np.abs(ref).max())

    dual = np.array([float(s) for s in open(tmp_path / "inst_dual.txt").read().splitlines()])
    assert np.array_equal(dual, np.concatenate([out.dual_source_solution, out.dual_target_solution]))

    summary = json.loads(open(tmp_path / "inst_summary.json").read())
    assert list(summary) == ["instance_name", "command_line_invocation", "termination_reason", "termination_string",
                             "iteration_count", "solve_time_sec", "solution_stats", "solution_type",
                             "iteration_stats"]
    assert summary["termination_reason"] == "TERMINATION_REASON_OPTIMAL"
    assert summary["solution_type"] == "POINT_TYPE_AVERAGE_ITERATE"
    assert summary["iteration_count"] == out.iteration_count
    assert summary["iteration_stats"] == []
    full = json.loads(gzip.open(tmp_path / "inst_full_log.json.gz", "rt").read())
    assert len(full["iteration_stats"]) == len(out.iteration_stats)
    last = full["iteration_stats"][-1]["convergence_information"][0]
    assert last["primal_objective"] == out.iteration_stats[-1].convergence_information[0].primal_objective
