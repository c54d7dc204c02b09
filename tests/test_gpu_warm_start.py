"""What the trained warm-start network (pdot_amd/warm_start_weights.npz,
written by python -m pdot_amd.train_warm_start) gains over the method
without it, on held-out Gaussian-mixture instances (seeds disjoint from the
training and evaluation seeds of the training script).

* Prediction: the predicted dual potentials are closer to the converged
  duals than the zero start is (mean squared error after the (u + t, v - t)
  normalisation), by at least 2x.
* Solve: the warm-started solve, with the KKT acceptance test and the zero-start
  fallback of learned.solve_warm, terminates OPTIMAL at the same objective
  as the zero start.  The iteration counts of both are printed: the
  restarted adaptive PDHG does not turn the better dual start into fewer
  iterations (it re-grows the primal from X0 = 0 while the duals move), so
  no iteration gain is asserted.
"""
import os

import numpy as np
import pytest

import pdot_amd as P
from pdot_amd import instances, learned

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
        warm, used, pred_sec = learned.solve_warm(P.default_parameters(tol, record_iteration_stats=False), p,
                                                  inst.x, inst.y, net)
        assert zero.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        assert warm.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        oz = zero.iteration_stats[-1].convergence_information[0].primal_objective
        ow = warm.iteration_stats[-1].convergence_information[0].primal_objective
        assert abs(oz - ow) <= 20 * tol * (1 + abs(oz))
        print(f"seed {seed}: zero {zero.iteration_count} it, warm {warm.iteration_count} it "
              f"(warm start used: {used}, predict {pred_sec * 1e3:.1f} ms)")
