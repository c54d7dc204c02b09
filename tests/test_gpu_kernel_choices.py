"""Kernel choices that must not change results.

* Where the controller runs (pdot_set_controller_fold): in the last block of
  the reduce kernel to finish, whose blocks publish their scalars with
  agent-scope stores and an agent-scope completion count instead of a
  device-scope fence, or in a one-wave kernel after the reduce kernel.  The
  controller sums the same scalars in the same order either way.
* How the step kernel streams the matrices (pdot_set_stream_policy): with
  the nontemporal hint or without it.  The arithmetic and its order are the
  same either way.

So whole solves must agree bit for bit: iterations, restarts, every
evaluation's step size and primal weight, the plan and the duals."""
import numpy as np
import pytest

import pdot_amd as P
from pdot_amd import _lib, instances

pytestmark = pytest.mark.gpu

CASES = [("dense", _lib.STORAGE_F64, lambda: instances.uniform_sqeuclid(300, 257, seed=21), 1e-6),
         ("dense", _lib.STORAGE_F32, lambda: instances.gaussian_shift(517, 260, seed=22), 1e-5),
         ("points", _lib.STORAGE_F32, lambda: instances.l1_mixture_3d_points(333, 210, seed=23), 1e-4),
         # 140 reduce blocks: within the size the default folds at
         ("dense", _lib.STORAGE_F64, lambda: instances.gaussian_shift(4500, 4400, seed=24), 1e-4)]


def _solve(prob, storage, points, params, fold=-1, stream=-1):
    n, m = len(prob.source_distribution), len(prob.target_distribution)
    dev = _lib.DeviceSolver(n, m, storage=storage)
    if points:
        dev.upload_points(prob)
    else:
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    dev.set_controller_fold(fold)
    dev.set_stream_policy(stream)
    out = P.optimize(params, prob, storage=storage, solver=dev)
    dev.close()
    return out


def _same_solves(outs):
    a = outs[0]
    assert a.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
    assert any(s.restart_used != 0 for s in a.iteration_stats)
    for b in outs[1:]:
        assert b.iteration_count == a.iteration_count
        assert len(b.iteration_stats) == len(a.iteration_stats)
        for sa, sb in zip(a.iteration_stats, b.iteration_stats):
            assert sa.restart_used == sb.restart_used
            assert sa.step_size == sb.step_size and sa.primal_weight == sb.primal_weight
            assert sa.convergence_information[0].primal_objective == sb.convergence_information[0].primal_objective
        assert np.array_equal(a.primal_solution, b.primal_solution)
        assert np.array_equal(a.dual_source_solution, b.dual_source_solution)
        assert np.array_equal(a.dual_target_solution, b.dual_target_solution)


@pytest.mark.parametrize("form,storage,make,tol", CASES)
def test_controller_placements_agree_bitwise(form, storage, make, tol):
    prob = make()
    params = P.default_parameters(tol)
    _same_solves([_solve(prob, storage, form == "points", params, fold=mode) for mode in (1, 0, -1)])


@pytest.mark.parametrize("form,storage,make,tol", CASES)
def test_stream_policies_agree_bitwise(form, storage, make, tol):
    prob = make()
    params = P.default_parameters(tol)
    _same_solves([_solve(prob, storage, form == "points", params, stream=mode) for mode in (1, 0, -1)])


def test_controller_placements_agree_with_constant_steps_and_a_warm_start():
    prob = instances.uniform_sqeuclid(200, 150, seed=25)
    g = np.random.default_rng(25)
    u0, v0 = 0.01 * g.standard_normal(200), 0.01 * g.standard_normal(150)
    from dataclasses import replace
    params = replace(P.default_parameters(1e-5), step_size_policy_params=P.ConstantStepsizeParams())
    outs = []
    for mode in (1, 0):
        dev = _lib.DeviceSolver(200, 150)
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
        dev.set_controller_fold(mode)
        outs.append(P.optimize(params, prob, solver=dev, warm_start=(u0, v0)))
        dev.close()
    a, b = outs
    assert a.iteration_count == b.iteration_count
    assert np.array_equal(a.primal_solution, b.primal_solution)
    assert np.array_equal(a.dual_source_solution, b.dual_source_solution)


def test_kernel_choice_modes_are_validated():
    dev = _lib.DeviceSolver(20, 30)
    with pytest.raises(_lib.PdotError, match="mode must be -1, 0 or 1"):
        dev.set_controller_fold(2)
    with pytest.raises(_lib.PdotError, match="mode must be -1, 0 or 1"):
        dev.set_stream_policy(-2)
    dev.close()
