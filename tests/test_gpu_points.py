"""GPU tests of the point-cloud cost form (pdot_upload_points): the cost is
formed in registers each pass from the point clouds instead of being read
from HBM.  The formula reproduces the dense generators' matrix bit for bit,
so in FP64 a point-cloud solve must equal the dense solve of that matrix bit
for bit (iterations, restarts, plan and duals); with FP32 storage (where the
dense form stores C rounded to FP32 and the point form keeps it to about one
FP64 ulp, multiplying by the reciprocal of the scale) results are pinned
against the HiGHS optima."""
import json
import os

import numpy as np
import pytest

import oracle as O
import pdot_amd as P
from pdot_amd import _lib, instances

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "lp_optima.json")))
POINTS = {"uniform_sqeuclid": instances.uniform_sqeuclid_points, "gaussian_shift": instances.gaussian_shift_points,
          "l1_mixture_3d": instances.l1_mixture_3d_points}


@pytest.mark.parametrize("make,args", [(instances.gaussian_shift_points, (700, 500, 3)),
                                       (instances.l1_mixture_3d_points, (400, 450, 4)),
                                       (instances.uniform_sqeuclid_points, (300, 257, 11)),
                                       (instances.uniform_sqeuclid_points, (1, 37, 2)),
                                       (instances.gaussian_shift_points, (41, 1, 6)),
                                       (instances.l1_mixture_3d_points, (1031, 17, 8))])
def test_point_solve_equals_dense_solve_bitwise(make, args):
    pts = make(*args)
    dense = pts.dense()
    params = P.default_parameters(1e-4)
    a = P.optimize(params, pts)
    b = P.optimize(params, dense)
    assert a.termination_reason == b.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
    assert a.iteration_count == b.iteration_count
    assert len(a.iteration_stats) == len(b.iteration_stats)
    for sa, sb in zip(a.iteration_stats, b.iteration_stats):
        assert sa.restart_used == sb.restart_used
        assert sa.convergence_information[0].primal_objective == sb.convergence_information[0].primal_objective
        assert sa.step_size == sb.step_size and sa.primal_weight == sb.primal_weight
    assert np.array_equal(a.primal_solution, b.primal_solution)
    assert np.array_equal(a.dual_source_solution, b.dual_source_solution)
    assert np.array_equal(a.dual_target_solution, b.dual_target_solution)


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_point_cached_ot_info_equals_dense(storage):
    pts = instances.gaussian_shift_points(3000, 2900, 5)  # two upload chunks
    dense = pts.dense()
    d1 = _lib.DeviceSolver(3000, 2900, storage=storage)
    d1.upload_points(pts)
    d2 = _lib.DeviceSolver(3000, 2900, storage=storage)
    d2.upload(dense.cost_matrix, dense.source_distribution, dense.target_distribution)
    assert d1.cached_ot_info() == d2.cached_ot_info()
    assert d1.cached_ot_info().l_inf_norm_primal_linear_objective == 1.0  # normalised by its maximum
    assert d1.device_bytes() < d2.device_bytes()  # no matrix allocated for C
    d1.close()
    d2.close()


@pytest.mark.parametrize("storage,tol", [(_lib.STORAGE_F32, 1e-4), (_lib.STORAGE_F64, 1e-6)])
def test_point_results_match_lp_optima(storage, tol):
    for c in FIX["lp"]:
        if c["generator"] not in POINTS or c["n"] * c["m"] > 1_000_000:
            continue
        pts = POINTS[c["generator"]](c["n"], c["m"], seed=c["seed"])
        out = P.optimize(P.default_parameters(tol), pts, storage=storage)
        assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL, c
        pobj = float(np.sum(pts.dense().cost_matrix * out.primal_solution))
        assert abs(pobj - c["optimum"]) <= 10 * tol * (1 + abs(c["optimum"])), c


def test_point_first_ten_evaluations_match_oracle():
    pts = instances.l1_mixture_3d_points(333, 270, 9)
    params = P.default_parameters(1e-4, iteration_limit=10)
    out = P.optimize(params, pts)
    ref = O.optimize(params, pts.dense(), mode=O.FUSED)
    assert out.iteration_count == ref["iteration_count"]
    for s_dev, s_ref in zip(out.iteration_stats, ref["stats"]):
        ci = s_dev.convergence_information[0]
        for f in ("primal_objective", "dual_objective", "l2_primal_residual", "l2_dual_residual"):
            np.testing.assert_allclose(getattr(ci, f), s_ref[f], rtol=1e-9, atol=1e-15, err_msg=f)


def test_point_form_beyond_2_31_elements_known_answer():
    """70000 x 31000 (2.17e9 elements, beyond 32-bit indexing), FP64 storage,
    point-cloud cost: the third evaluation reports ||X2||_F and <C, X2> of
    the closed-form plan after pass 2 (tests/point_kat.py), to 1e-11; the
    feasibility rounding of the returned plan meets the masses exactly."""
    from point_kat import third_evaluation
    n, m = 70000, 31000
    assert n * m > 2 ** 31
    g = np.random.default_rng(11)
    a, b = g.random(n), g.random(m)  # balanced masses (each sums to 1), as the rounding requires
    pts = P.PointCloudProblem(g.normal(size=(n, 2)), g.normal(size=(m, 2)) + [2.0, 0.0],
                              a / a.sum(), b / b.sum(), "sqeuclid", normalize=False)
    dev = _lib.DeviceSolver(n, m, storage=_lib.STORAGE_F64)
    dev.upload_points(pts)
    out = P.optimize(P.default_parameters(1e-4, iteration_limit=3), pts, storage=_lib.STORAGE_F64, solver=dev)
    assert out.iteration_count == 2 and len(out.iteration_stats) == 3
    ci = out.iteration_stats[2].convergence_information[0]
    norm, obj = third_evaluation(pts, block=256)
    print(f"||X2||_F {ci.l2_primal_variable:.15g} (host {norm:.15g}), <C,X2> {ci.primal_objective:.15g} (host {obj:.15g})")
    assert ci.l2_primal_variable == pytest.approx(norm, rel=1e-11)
    assert ci.primal_objective == pytest.approx(obj, rel=1e-11)
    # feasibility rounding of the returned plan at the same size: rows and columns sum to the masses
    Y = dev.round(dev.output_point)
    dev.close()
    del out
    np.testing.assert_allclose(Y.sum(axis=1), pts.source_distribution, rtol=0, atol=1e-13)
    np.testing.assert_allclose(Y.sum(axis=0), pts.target_distribution, rtol=0, atol=1e-13)
    assert Y.min() >= -1e-15  # er_i = a_i - rowsum_i can round to a tiny negative, as in the reference
