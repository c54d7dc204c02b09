"""What the learned starting state (pdot_amd/warm_start_weights.npz, written
by python -m pdot_amd.train_warm_start: the dual network and the calibrated
initial primal weight) gains over the method without it, on held-out
Gaussian-mixture instances (seeds disjoint from the training, calibration
and evaluation seeds of the training script).

* Prediction: the predicted dual potentials are closer to the converged
  duals than the zero start is (mean squared error after the (u + t, v - t)
  normalisation), by at least 2x.
* Solve: learned.solve_warm (KKT acceptance test for the duals, calibrated
  initial primal weight, plain zero-start fallback) terminates OPTIMAL at
  the zero start's objective to the tolerance.
* Gain: on the eight held-out 4096 x 4096 instances of the benchmark, the
  learned start needs at least 2x fewer iterations in total than the zero
  start.  The gain comes from the initial primal weight: the restarted
  adaptive PDHG does not turn the better dual start into fewer iterations
  (it re-grows the primal from X0 = 0 while the duals move), and the
  acceptance test rejects the predicted duals on these instances.
"""
import os

import numpy as np
import pytest

import pdot_amd as P
from pdot_amd import _lib, instances, learned

pytestmark = pytest.mark.gpu

HELD_OUT = [900100 + k for k in range(3)]


@pytest.fixture(scope="module")
def net():
    assert os.path.exists(learned.DEFAULT_WEIGHTS), "trained weights are missing"
    n = learned.WarmStartNet(learned.load_params())
    yield n
    n.close()


def test_network_predicts_duals_better_than_zero(net):
    ratios = []
    for seed in HELD_OUT:
        inst = instances.gmm_dirichlet(512, seed=seed)
        p = inst.problem
        out = P.optimize(P.default_parameters(1e-6, record_iteration_stats=False), p)
        assert out.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        a, b = p.source_distribution, p.target_distribution
        tu, tv = learned.normalize_duals(out.dual_source_solution, out.dual_target_solution, a, b)
        pu, pv = learned.normalize_duals(*net.predict_duals(inst.x, inst.y, a, b), a, b)
        target = np.concatenate([tu, tv])
        mse_pred = np.mean((np.concatenate([pu, pv]) - target) ** 2)
        mse_zero = np.mean(target ** 2)
        ratios.append(mse_pred / mse_zero)
    print("held-out prediction MSE / zero-start MSE:", [f"{r:.3f}" for r in ratios])
    assert max(ratios) < 0.5


def test_warm_started_solve_is_checked_and_correct(net):
    tol = 1e-4
    for seed in HELD_OUT:
        inst = instances.gmm_dirichlet(512, seed=seed)
        p = inst.problem
        zero = P.optimize(P.default_parameters(tol, record_iteration_stats=False), p)
        warm, info = learned.solve_warm(P.default_parameters(tol, record_iteration_stats=False), p,
                                        inst.x, inst.y, net)
        assert zero.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        assert warm.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        oz = zero.iteration_stats[-1].convergence_information[0].primal_objective
        ow = warm.iteration_stats[-1].convergence_information[0].primal_objective
        assert abs(oz - ow) <= 20 * tol * (1 + abs(oz))
        print(f"seed {seed}: zero {zero.iteration_count} it, warm {warm.iteration_count} it "
              f"(duals used: {info.used_duals}, primal weight {info.primal_weight}, "
              f"predict {info.predict_sec * 1e3:.1f} ms)")


def test_learned_start_halves_iterations_on_held_out_4096(net):
    tol = 1e-4
    total_zero = total_warm = 0
    for seed in [900200 + k for k in range(8)]:
        inst = instances.gmm_dirichlet(4096, seed=seed)
        p = inst.problem
        prm = P.default_parameters(tol, record_iteration_stats=False)
        zero = P.optimize(prm, p, storage=_lib.STORAGE_F32)
        warm, info = learned.solve_warm(prm, p, inst.x, inst.y, net, storage=_lib.STORAGE_F32)
        assert zero.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        assert warm.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        assert not info.fell_back
        oz = zero.iteration_stats[-1].convergence_information[0].primal_objective
        ow = warm.iteration_stats[-1].convergence_information[0].primal_objective
        assert abs(oz - ow) <= 20 * tol * (1 + abs(oz))
        total_zero += zero.iteration_count
        total_warm += warm.iteration_count
        print(f"seed {seed}: zero {zero.iteration_count} it, learned start {warm.iteration_count} it "
              f"(duals used: {info.used_duals}, primal weight {info.primal_weight})")
    print(f"total: zero {total_zero} it, learned start {total_warm} it ({total_zero / total_warm:.2f}x)")
    assert total_zero >= 2 * total_warm


def test_primal_weight_calibration_picks_the_fewest_iterations():
    """calibrate_primal_weight on two small calibration instances: the table
    holds one iteration count per instance and weight, and the weight it
    returns has the smallest total."""
    from pdot_amd import train_warm_start as T
    best, table = T.calibrate_primal_weight(count=2, n=384, grid=(0.2, 1.0), log=lambda s: None)
    assert set(table) == {0.2, 1.0}
    assert all(len(v) == 2 and all(it > 0 for it in v) for v in table.values())
    assert sum(table[best]) == min(sum(v) for v in table.values())
