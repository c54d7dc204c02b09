"""Full-size GPU checks (10k x 10k Gaussian shift, 30k x 30k L1 mixture) by a
certificate that needs no oracle run: for any duals u,

    LB(u) = sum_i a_i u_i + sum_j b_j min_i (C_ij - u_i)

is the objective of a feasible dual point (the c-transform), so LB <= OPT;
and the feasibility rounding Y of the returned plan is a feasible plan, so
<C, Y> = UB >= OPT.  The certified gap is bounded by the solver's own final
residuals (absolute L2 primal residual P2, dual residual D2, objectives):

    UB - LB <= |pobj - dobj| + ||b||_2 D2 + 2 ||C||_inf sqrt(n + m) P2

(the c-transform loses at most b'|min(G, 0)| <= ||b|| D2 against the dual
objective; the rounding moves the plan by at most 2 ||r(X) - a, c(X) - b||_1
in l1, Altschuler-Weed-Rigollet), and |pobj - UB| by the last term alone.
Y must be feasible to rounding.  FP32 storage solves the FP32-rounded
problem while the bounds use the FP64 cost: 1e-6 relative slack."""
import numpy as np
import pytest

import pdot_amd as P
from pdot_amd import _lib, instances

pytestmark = pytest.mark.gpu


def _certificate(C_cols, a, b, u, Y_cols):
    """LB(u) and UB = <C, Y>, streaming the columns of C and Y in chunks
    (C_cols(j0, j1) returns C[:, j0:j1])."""
    m = len(b)
    cmin = np.empty(m)
    ub = 0.0
    for j0 in range(0, m, 1000):
        j1 = min(m, j0 + 1000)
        Cc = C_cols(j0, j1)
        cmin[j0:j1] = (Cc - u[:, None]).min(axis=0)
        ub += float(np.sum(Cc * Y_cols(j0, j1)))
    return float(a @ u + b @ cmin), ub


def _check_certificate(out, lb, ub, a, b, c_inf, storage):
    ci = out.iteration_stats[-1].convergence_information[0]
    n, m = len(a), len(b)
    rounding_bound = 2.0 * c_inf * np.sqrt(n + m) * ci.l2_primal_residual
    gap_bound = abs(ci.primal_objective - ci.dual_objective) + np.linalg.norm(b) * ci.l2_dual_residual + rounding_bound
    slack = 1e-12 + (1e-6 * abs(ub) if storage == _lib.STORAGE_F32 else 0.0)
    print(f"{n}x{m} storage {storage}: LB {lb:.10g} UB {ub:.10g} certified gap {(ub - lb) / abs(ub):.3e} "
          f"(bound {gap_bound / abs(ub):.3e}); reported objective {ci.primal_objective:.10g}")
    assert lb <= ub + slack
    assert ub - lb <= gap_bound + slack
    assert abs(ci.primal_objective - ub) <= rounding_bound + slack


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_gaussian_shift_10k_certified_gap(storage):
    prob = instances.gaussian_shift(10000)
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    dev = _lib.DeviceSolver(10000, 10000, storage=storage)
    dev.upload(C, a, b)
    out = P.optimize(P.default_parameters(1e-4, record_iteration_stats=False), prob, storage=storage, solver=dev)
    assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
    Y = dev.round(dev.output_point)
    dev.close()
    np.testing.assert_allclose(Y.sum(axis=1), a, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Y.sum(axis=0), b, rtol=0, atol=1e-12)
    assert Y.min() >= -1e-15
    lb, ub = _certificate(lambda j0, j1: C[:, j0:j1], a, b, out.dual_source_solution, lambda j0, j1: Y[:, j0:j1])
    _check_certificate(out, lb, ub, a, b, float(np.abs(C).max()), storage)


def test_l1_mixture_30k_fp32_certified_gap():
    pts = instances.l1_mixture_3d_points(30000)
    x, y, a, b = pts.source_points, pts.target_points, pts.source_distribution, pts.target_distribution
    dev = _lib.DeviceSolver(30000, 30000, storage=_lib.STORAGE_F32)
    dev.upload_points(pts)  # the cost formed on the device; no 7.2 GB matrix on the host
    out = P.optimize(P.default_parameters(1e-4, record_iteration_stats=False), pts, storage=_lib.STORAGE_F32,
                     solver=dev)
    assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
    Y = dev.round(dev.output_point)
    dev.close()
    np.testing.assert_allclose(Y.sum(axis=1), a, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Y.sum(axis=0), b, rtol=0, atol=1e-12)

    def cols(j0, j1):  # the L1 cost, column block, as the dense generator forms it
        ct = np.zeros((j1 - j0, x.shape[0]))
        for d in range(x.shape[1]):
            ct += np.abs(y[j0:j1, d][:, None] - x[:, d][None, :])
        return ct.T

    lb, ub = _certificate(cols, a, b, out.dual_source_solution, lambda j0, j1: Y[:, j0:j1])
    cinf = max(float(cols(j0, min(30000, j0 + 1000)).max()) for j0 in range(0, 30000, 1000))
    _check_certificate(out, lb, ub, a, b, cinf, _lib.STORAGE_F32)


def test_l1_mixture_30k_dense_fp32_certified_gap():
    """The configured form: the dense 30000 x 30000 L1 cost matrix (7.2 GB in
    FP64 on the host) uploaded and narrowed to FP32 on the device."""
    prob = instances.l1_mixture_3d(30000)
    C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
    dev = _lib.DeviceSolver(30000, 30000, storage=_lib.STORAGE_F32)
    dev.upload(C, a, b)
    out = P.optimize(P.default_parameters(1e-4, record_iteration_stats=False), prob, storage=_lib.STORAGE_F32,
                     solver=dev)
    assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
    Y = dev.round(dev.output_point)
    dev.close()
    np.testing.assert_allclose(Y.sum(axis=1), a, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Y.sum(axis=0), b, rtol=0, atol=1e-12)
    lb, ub = _certificate(lambda j0, j1: C[:, j0:j1], a, b, out.dual_source_solution, lambda j0, j1: Y[:, j0:j1])
    _check_certificate(out, lb, ub, a, b, float(np.abs(C).max()), _lib.STORAGE_F32)


def test_batched_uniform_instances_certified():
    """All 512 instances of the batch workload (n, m in [64, 2048], FP32
    storage) solved together with the calibrated initial primal weight 0.2
    (the reference's default 1.0 leaves a tail of 10^5-10^6-iteration
    instances): every instance reaches OPTIMAL and carries the certificate
    (rounding of the downloaded plan on the device)."""
    from dataclasses import replace
    probs = instances.batched_uniform(512)
    from pdot_amd.batch import optimize_batch
    params = replace(P.default_parameters(1e-4, iteration_limit=100000, record_iteration_stats=False),
                     primal_importance=0.2)
    outs = optimize_batch(params, probs, storage=_lib.STORAGE_F32)
    for prob, out in zip(probs, outs):
        assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        C, a, b = prob.cost_matrix, prob.source_distribution, prob.target_distribution
        Y = P.rounding(prob, out.primal_solution.copy(order="F"))
        lb, ub = _certificate(lambda j0, j1: C[:, j0:j1], a, b, out.dual_source_solution, lambda j0, j1: Y[:, j0:j1])
        _check_certificate(out, lb, ub, a, b, float(np.abs(C).max()), _lib.STORAGE_F32)
