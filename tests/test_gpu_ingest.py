"""GPU tests of the ingest path (pdot_upload, pdot_cached_ot_info): the
pinned, chunked upload of ot_cpu_to_gpu (src/io.jl:18-36) with the FP32
narrowing on the device, and cached_ot_info (src/termination.jl:94-101)
computed in the same pass, against the oracle."""
import math

import numpy as np
import pytest

import oracle as O
import pdot_amd as P
from pdot_amd import _lib, instances

pytestmark = pytest.mark.gpu

EVAL_FIELDS = ("primal_objective", "dual_objective", "l2_primal_residual", "l2_dual_residual",
               "l_inf_dual_residual", "l2_primal_variable")


def _info_tuple(info):
    return (info.l_inf_norm_primal_linear_objective, info.l_inf_norm_primal_right_hand_side,
            info.l2_norm_primal_linear_objective, info.l2_norm_primal_right_hand_side)


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_multi_chunk_upload_matches_oracle(storage):
    """3000 x 2900 FP64 is 69.6 MB: two staging chunks of whole columns.  The
    norms agree with the oracle (maxima exactly), and the first ten
    evaluations of a solve on the uploaded matrix match the oracle's."""
    prob = instances.uniform_sqeuclid(3000, 2900, seed=31)
    dev = _lib.DeviceSolver(3000, 2900, storage=storage)
    dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    got = _info_tuple(dev.cached_ot_info())
    ref = O.cached_ot_info(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    assert got[0] == ref[0] and got[1] == ref[1]
    np.testing.assert_allclose(got[2:], ref[2:], rtol=1e-13)
    params = P.default_parameters(1e-4, iteration_limit=10)
    out = P.optimize(params, prob, storage=storage, solver=dev)
    r = O.optimize(params, prob, mode=O.FUSED)
    assert out.iteration_count == r["iteration_count"]
    if storage == _lib.STORAGE_F64:
        for s_dev, s_ref in zip(out.iteration_stats, r["stats"]):
            ci = s_dev.convergence_information[0]
            for f in EVAL_FIELDS:
                np.testing.assert_allclose(getattr(ci, f), s_ref[f], rtol=1e-9, atol=1e-15, err_msg=f)
    else:
        # FP32 storage computes on the narrowed matrix: agreement to FP32
        # rounding (a misplaced chunk would be off by O(1))
        for s_dev, s_ref in zip(out.iteration_stats, r["stats"]):
            ci = s_dev.convergence_information[0]
            for f in ("primal_objective", "dual_objective"):
                np.testing.assert_allclose(getattr(ci, f), s_ref[f], rtol=1e-4, atol=1e-9, err_msg=f)
    dev.close()


def test_cached_ot_info_edge_cases():
    rng = np.random.default_rng(3)
    a, b = np.full(70, 1 / 70), np.full(50, 1 / 50)
    cases = {"zero": np.zeros((70, 50)), "negative max": rng.random((70, 50)), "nan": rng.random((70, 50))}
    cases["negative max"][13, 7] = -5.0
    cases["nan"][2, 40] = math.nan
    for name, C in cases.items():
        dev = _lib.DeviceSolver(70, 50)
        dev.upload(np.asfortranarray(C), a, b)
        got = _info_tuple(dev.cached_ot_info())
        ref = O.cached_ot_info(C, a, b)
        if name == "nan":
            assert math.isnan(got[0]) and math.isnan(got[2]), name
        else:
            assert got[0] == ref[0], name
            np.testing.assert_allclose(got[2], ref[2], rtol=1e-14, err_msg=name)
        assert got[1] == ref[1] and got[3] == pytest.approx(ref[3], rel=1e-15)
        dev.close()


def test_reupload_replaces_problem():
    """A second upload into the same context replaces C, a, b and the norms."""
    p1 = instances.uniform_sqeuclid(200, 150, seed=1)
    p2 = instances.uniform_sqeuclid(200, 150, seed=2)
    dev = _lib.DeviceSolver(200, 150)
    dev.upload(p1.cost_matrix, p1.source_distribution, p1.target_distribution)
    dev.upload(p2.cost_matrix, p2.source_distribution, p2.target_distribution)
    ref = O.cached_ot_info(p2.cost_matrix, p2.source_distribution, p2.target_distribution)
    assert _info_tuple(dev.cached_ot_info())[0] == ref[0]
    params = P.default_parameters(1e-4, iteration_limit=10)
    out = P.optimize(params, p2, solver=dev)
    r = O.optimize(params, p2, mode=O.FUSED)
    np.testing.assert_allclose(out.iteration_stats[-1].convergence_information[0].primal_objective,
                               r["stats"][-1]["primal_objective"], rtol=1e-9)
    dev.close()
