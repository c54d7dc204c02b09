"""Deferred averaging (on by default for single problems): an ordinary pass
may leave its contribution to the running average (X, u, v, r, c) to the
next pass, which applies it to the average it loads, rounded to the storage
type as a store would round it, before adding its own.  Every second pass
then neither reads nor writes the average matrix.  Each element's arithmetic
is unchanged, so every result must equal the one with deferral off bit for
bit: whole solves (iterations, restarts, evaluations, plan and duals), and a
stepping sequence that reads the average while a contribution lags
(downloads, evaluations, rounding, restarts, deferral switched mid-solve)."""
import numpy as np
import pytest

import pdot_amd as P
from pdot_amd import _lib, instances

pytestmark = pytest.mark.gpu

# the solver state's numeric fields; the slot indices and the deferral
# bookkeeping legitimately differ between the two paths
_STATE_FIELDS = ("step_size", "primal_weight", "cumulative_kkt_passes", "sum_weights", "total_number_iterations",
                 "avg_count", "take_step_calls", "error_call", "last_interaction", "last_movement", "last_step_limit",
                 "last_dx2", "last_dot", "last_dy2", "numerical_error", "gzero", "pending")


def _same_state(s1, s2):
    for k in _STATE_FIELDS:
        assert getattr(s1, k) == getattr(s2, k), k


def _same_arrays(xs, ys):
    for x, y in zip(xs, ys):
        assert np.array_equal(x, y)


def _device(prob, storage, defer, points):
    n = len(prob.source_distribution)
    m = len(prob.target_distribution)
    dev = _lib.DeviceSolver(n, m, storage=storage)
    if points:
        dev.upload_points(prob)
    else:
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    dev.set_defer(defer)
    return dev


CASES = [("dense", _lib.STORAGE_F64, lambda: instances.gaussian_shift(300, 257, seed=11)),
         ("dense", _lib.STORAGE_F32, lambda: instances.gaussian_shift(300, 257, seed=11)),
         ("dense", _lib.STORAGE_F32, lambda: instances.uniform_sqeuclid(61, 1003, seed=12)),
         ("points", _lib.STORAGE_F64, lambda: instances.l1_mixture_3d_points(400, 450, seed=13)),
         ("points", _lib.STORAGE_F32, lambda: instances.gaussian_shift_points(517, 260, seed=14))]


@pytest.mark.parametrize("form,storage,make", CASES)
def test_deferred_solve_equals_eager_bitwise(form, storage, make):
    prob = make()
    outs = []
    for defer in (True, False):
        dev = _device(prob, storage, defer, form == "points")
        out = P.optimize(P.default_parameters(1e-4, iteration_limit=400000), prob, storage=storage, solver=dev)
        avg = dev.download(_lib.POINT_AVERAGE)
        cur = dev.download(_lib.POINT_CURRENT)
        Y = dev.round(_lib.POINT_AVERAGE)
        dev.close()
        outs.append((out, avg, cur, Y))
    (a, avg_a, cur_a, Ya), (b, avg_b, cur_b, Yb) = outs
    assert a.termination_reason == b.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
    assert a.iteration_count == b.iteration_count > 100
    assert len(a.iteration_stats) == len(b.iteration_stats)
    for sa, sb in zip(a.iteration_stats, b.iteration_stats):
        assert sa.restart_used == sb.restart_used
        assert sa.step_size == sb.step_size and sa.primal_weight == sb.primal_weight
        for ca, cb in zip(sa.convergence_information, sb.convergence_information):
            assert ca.primal_objective == cb.primal_objective and ca.dual_objective == cb.dual_objective
            assert ca.l2_primal_residual == cb.l2_primal_residual and ca.l2_dual_residual == cb.l2_dual_residual
    _same_arrays((a.primal_solution, a.dual_source_solution, a.dual_target_solution),
                 (b.primal_solution, b.dual_source_solution, b.dual_target_solution))
    _same_arrays(avg_a + cur_a + (Ya,), avg_b + cur_b + (Yb,))


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_lagging_average_reads_equal_eager_updates(storage):
    prob = instances.gaussian_shift(200, 170, seed=7)
    devs = [_device(prob, storage, defer, False) for defer in (True, False)]

    def both(f):
        return [f(d) for d in devs]

    def steps(k):
        _same_state(*both(lambda d: d.run_steps(k)))

    def lag():  # single passes until the deferring solver holds a lagging contribution
        for _ in range(2):
            if devs[0].get_state().avg_pending:
                break
            steps(1)
        assert devs[0].get_state().avg_pending == 1 and devs[1].get_state().avg_pending == 0

    def check_average(average=True):
        if average:
            _same_arrays(*both(lambda d: d.download(_lib.POINT_AVERAGE)))
        _same_arrays(*both(lambda d: d.download(_lib.POINT_CURRENT)))
        _same_state(*both(lambda d: d.get_state()))

    def check_eval():
        evs = both(lambda d: d.eval())
        assert evs[0][0].as_dict() == evs[1][0].as_dict() and evs[0][1].as_dict() == evs[1][1].as_dict()
        _same_state(evs[0][2], evs[1][2])

    both(lambda d: d.set_start(primal_weight=1.5))
    steps(4)
    lag()
    check_average()  # downloads materialize the lagging contribution
    steps(5)
    lag()
    check_eval()  # an evaluation pass applies the lagging contribution
    steps(4)  # commits the evaluation's pass first
    lag()
    _same_arrays(*both(lambda d: [d.round(_lib.POINT_AVERAGE)]))
    steps(7)
    lag()
    both(lambda d: d.restart(True, 2.25))  # restart to an average whose last contribution lags
    check_average(average=False)  # the restart empties the average
    steps(1)
    check_average()
    steps(9)
    lag()
    devs[0].set_defer(False)  # switched off with a contribution lagging, then on again
    steps(3)
    devs[0].set_defer(True)
    steps(11)
    lag()
    check_eval()
    both(lambda d: d.discard())  # a discarded evaluation leaves the lag in place
    assert devs[0].get_state().avg_pending == 1
    check_average()
    steps(2)
    lag()
    both(lambda d: d.restart(False, 1.0))  # restart to the current point drops the lag
    check_average(average=False)
    steps(5)
    check_average()
    both(lambda d: d.close())


def _batch(probs, storage, defer):
    b = _lib.BatchSolver([p.num_row for p in probs], [p.num_col for p in probs], storage=storage)
    for k, p in enumerate(probs):
        b.upload(k, p.cost_matrix, p.source_distribution, p.target_distribution)
    b.set_defer(defer)
    return b


BATCH_SIZES = [(64, 64), (300, 257), (1, 40), (50, 1), (1100, 70), (130, 1500)]


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_batch_deferred_solve_equals_eager_bitwise(storage):
    from pdot_amd.batch import optimize_batch
    probs = [instances.uniform_sqeuclid(n, m, seed=20 + k) for k, (n, m) in enumerate(BATCH_SIZES)]
    runs = []
    for defer in (True, False):
        b = _batch(probs, storage, defer)
        outs = optimize_batch(P.default_parameters(1e-4, iteration_limit=400000), probs, storage=storage, batch=b)
        states = b.get_states()
        avgs = [b.download(k, _lib.POINT_AVERAGE) if states[k].avg_count > 0 else None for k in range(len(probs))]
        b.close()
        runs.append((outs, avgs))
    (oa, aa), (ob, ab) = runs
    for a, b in zip(oa, ob):
        assert a.termination_reason == b.termination_reason
        assert a.iteration_count == b.iteration_count
        assert len(a.iteration_stats) == len(b.iteration_stats)
        for sa, sb in zip(a.iteration_stats, b.iteration_stats):
            assert sa.restart_used == sb.restart_used and sa.step_size == sb.step_size
            ca, cb = sa.convergence_information[0], sb.convergence_information[0]
            assert ca.primal_objective == cb.primal_objective and ca.l2_dual_residual == cb.l2_dual_residual
        _same_arrays((a.primal_solution, a.dual_source_solution, a.dual_target_solution),
                     (b.primal_solution, b.dual_source_solution, b.dual_target_solution))
    assert sum(a.iteration_count for a in oa) > 500
    for x, y in zip(aa, ab):
        assert (x is None) == (y is None)
        if x is not None:
            _same_arrays(x, y)


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_batch_lagging_average_reads_equal_eager_updates(storage):
    probs = [instances.uniform_sqeuclid(n, m, seed=40 + k) for k, (n, m) in enumerate([(70, 90), (257, 33), (5, 300)])]
    bs = [_batch(probs, storage, defer) for defer in (True, False)]
    K = len(probs)

    def both(f):
        return [f(b) for b in bs]

    def same_states(s1, s2):
        for x, y in zip(s1, s2):
            _same_state(x, y)

    def check(average=(True,) * K):
        same_states(*both(lambda b: b.get_states()))
        for k in range(K):
            _same_arrays(*both(lambda b: b.download(k, _lib.POINT_CURRENT)))
            if average[k]:
                _same_arrays(*both(lambda b: b.download(k, _lib.POINT_AVERAGE)))
        same_states(*both(lambda b: b.get_states()))

    def check_eval():
        evs = both(lambda b: b.eval())
        for k in range(K):
            assert evs[0][0][k].as_dict() == evs[1][0][k].as_dict()
            assert evs[0][1][k].as_dict() == evs[1][1][k].as_dict()
        same_states(evs[0][2], evs[1][2])

    both(lambda b: b.set_start(_lib.POLICY_ADAPTIVE, 2.0, 1.0, 1.0, [1.0, 0.5, 2.0]))
    same_states(*both(lambda b: b.run_steps([4, 5, 6])))  # instances 0 and 2 end with a lagging contribution
    assert [s.avg_pending for s in bs[0].get_states()] == [1, 0, 1]
    assert [s.avg_pending for s in bs[1].get_states()] == [0, 0, 0]
    check()
    same_states(*both(lambda b: b.run_steps([3, 4, 3])))
    check_eval()  # evaluation passes that apply lagging contributions
    same_states(*both(lambda b: b.run_steps([2, 2, 2])))  # commits the evaluations' passes first
    same_states(*both(lambda b: b.run_steps([1, 2, 1])))
    pend = [s.avg_pending for s in bs[0].get_states()]
    assert any(pend)
    both(lambda b: b.restart([2, 2, 1], [1.5, 0.75, 3.0]))  # to the average (some lagging) and to the current
    check(average=(False,) * K)
    same_states(*both(lambda b: b.run_steps([6, 7, 8])))
    bs[0].set_defer(False)  # switched off with contributions lagging, then on again
    same_states(*both(lambda b: b.run_steps([3, 3, 3])))
    bs[0].set_defer(True)
    same_states(*both(lambda b: b.run_steps([9, 10, 11])))
    check_eval()
    both(lambda b: b.discard([1, 1, 1]))
    check()
    both(lambda b: b.close())


def _policy_params(kind):
    p = P.default_parameters(1e-4, iteration_limit=6000)
    if kind == "constant_step":
        p.step_size_policy_params = P.ConstantStepsizeParams()
    elif kind == "fixed_frequency":
        p.restart_params = P.construct_restart_parameters(P.FIXED_FREQUENCY, P.KKT_GREEDY, 256, 0.9, 0.1, 0.9, 0.5)
    elif kind == "no_restarts":
        p.restart_params = P.construct_restart_parameters(P.NO_RESTARTS, P.KKT_GREEDY, 1000, 0.36, 0.1, 0.9, 0.5)
    elif kind == "no_restart_to_current":
        p.restart_params = P.construct_restart_parameters(P.ADAPTIVE_KKT, P.NO_RESTART_TO_CURRENT, 1000, 0.36, 0.1,
                                                          0.9, 0.5)
    return p


@pytest.mark.parametrize("kind", ["constant_step", "fixed_frequency", "no_restarts", "no_restart_to_current",
                                  "warm_start"])
def test_deferred_policy_variants_equal_eager_bitwise(kind):
    """The step-size and restart policies change when the average is reset,
    restarted to or read; the deferred and eager runs must still agree bit
    for bit (also from a warm start, where G starts from the given duals)."""
    prob = instances.uniform_sqeuclid(90, 70, seed=31)
    ws = None
    if kind == "warm_start":
        rng = np.random.default_rng(3)
        ws = (0.01 * rng.standard_normal(90), 0.01 * rng.standard_normal(70))
    outs = []
    for defer in (True, False):
        dev = _device(prob, _lib.STORAGE_F64, defer, False)
        outs.append(P.optimize(_policy_params(kind), prob, solver=dev, warm_start=ws))
        dev.close()
    a, b = outs
    assert a.termination_reason == b.termination_reason
    assert a.iteration_count == b.iteration_count > 100
    assert len(a.iteration_stats) == len(b.iteration_stats)
    for sa, sb in zip(a.iteration_stats, b.iteration_stats):
        assert sa.restart_used == sb.restart_used
        assert sa.step_size == sb.step_size and sa.primal_weight == sb.primal_weight
        for ca, cb in zip(sa.convergence_information, sb.convergence_information):
            assert ca.primal_objective == cb.primal_objective and ca.l2_dual_residual == cb.l2_dual_residual
    _same_arrays((a.primal_solution, a.dual_source_solution, a.dual_target_solution),
                 (b.primal_solution, b.dual_source_solution, b.dual_target_solution))


PASS_KERNEL_CASES = CASES + [("dense", _lib.STORAGE_F64, lambda: instances.l1_mixture_3d(210, 333, seed=15)),
                             ("points", _lib.STORAGE_F32, lambda: instances.l1_mixture_3d_points(333, 210, seed=16))]


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("form,storage,make", PASS_KERNEL_CASES)
def test_pass_kernels_equal_run_time_kernel_bitwise(form, storage, make, graphs):
    """The kernels compiled for the passes that leave the average alone
    (chosen by the host from the alternation of the pass kinds, also inside
    captured graphs) give exactly the results of the kernel that decides per
    pass at run time: whole solves under restarts, a warm start and a
    constant step size, and the final points."""
    from dataclasses import replace
    prob = make()
    n = len(prob.source_distribution)
    rng = np.random.default_rng(n)
    warm = (0.01 * rng.standard_normal(n), 0.01 * rng.standard_normal(len(prob.target_distribution)))
    base = P.default_parameters(1e-4, iteration_limit=400000)
    runs = [("default", base, None), ("warm start", base, warm),
            ("constant step", replace(base, step_size_policy_params=P.ConstantStepsizeParams()), None)]
    for name, params, ws in runs:
        outs = []
        for pass_kernels in (True, False):
            dev = _device(prob, storage, True, form == "points")
            dev.set_pass_kernels(pass_kernels)
            out = P.optimize(params, prob, storage=storage, solver=dev, warm_start=ws, use_graphs=graphs)
            outs.append((out, dev.download(_lib.POINT_AVERAGE), dev.download(_lib.POINT_CURRENT), dev.get_state()))
            dev.close()
        (a, avg_a, cur_a, st_a), (b, avg_b, cur_b, st_b) = outs
        assert a.termination_reason == b.termination_reason, name
        assert a.iteration_count == b.iteration_count > 10, name
        assert [s.restart_used for s in a.iteration_stats] == [s.restart_used for s in b.iteration_stats], name
        for sa, sb in zip(a.iteration_stats, b.iteration_stats):
            assert sa.step_size == sb.step_size and sa.primal_weight == sb.primal_weight, name
            ca, cb = sa.convergence_information[0], sb.convergence_information[0]
            assert ca.primal_objective == cb.primal_objective and ca.l2_dual_residual == cb.l2_dual_residual, name
        _same_arrays((a.primal_solution, a.dual_source_solution, a.dual_target_solution),
                     (b.primal_solution, b.dual_source_solution, b.dual_target_solution))
        _same_arrays(avg_a + cur_a, avg_b + cur_b)
        _same_state(st_a, st_b)
        assert st_a.numerical_error == 0
