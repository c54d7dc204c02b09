"""GPU tests of the C ABI's error behaviour: calls out of order or with bad
arguments return a PDOT_E_* code with a message (raised as PdotError by the
binding, as the reference raises a Julia error), and leave the context usable."""
import numpy as np
import pytest

import pdot_amd as P
from pdot_amd import _lib, instances

pytestmark = pytest.mark.gpu


def test_state_errors_are_reported_and_recoverable():
    prob = instances.uniform_sqeuclid(50, 40, seed=3)
    dev = _lib.DeviceSolver(50, 40)
    with pytest.raises(_lib.PdotError, match="before pdot_upload"):
        dev.cached_ot_info()
    with pytest.raises(_lib.PdotError, match="before pdot_upload"):
        dev.set_start()
    dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    for call in (lambda: dev.run_steps(3), lambda: dev.eval(), lambda: dev.download(),
                 lambda: dev.round(), lambda: dev.time_steps(2)):
        with pytest.raises(_lib.PdotError, match="before pdot_set_start"):
            call()
    dev.set_start()
    with pytest.raises(_lib.PdotError, match="average is empty"):
        dev.download(_lib.POINT_AVERAGE)
    with pytest.raises(_lib.PdotError, match="average is empty"):
        dev.round(_lib.POINT_AVERAGE)
    # still usable: a solve on the same context matches a fresh one
    out = P.optimize(P.default_parameters(1e-4), prob, solver=dev)
    ref = P.optimize(P.default_parameters(1e-4), prob)
    assert out.iteration_count == ref.iteration_count
    assert np.array_equal(out.primal_solution, ref.primal_solution)
    dev.close()


def test_argument_errors():
    dev = _lib.DeviceSolver(20, 30)
    with pytest.raises(ValueError):
        dev.upload(np.zeros((30, 20)), np.ones(20) / 20, np.ones(30) / 30)
    bad = P.PointCloudProblem(np.zeros((20, 2)), np.zeros((30, 2)), np.ones(20) / 20, np.ones(30) / 30)
    with pytest.raises(_lib.PdotError, match="cannot normalise a zero cost"):
        dev.upload_points(P.PointCloudProblem(bad.source_points, bad.target_points, bad.source_distribution,
                                              bad.target_distribution, normalize=True))
    with pytest.raises(_lib.PdotError, match="unknown point"):
        dev.upload_points(bad)
        dev.set_start()
        dev.download(7)
    dev.close()
    with pytest.raises(_lib.PdotError):
        _lib.DeviceSolver(0, 10)


@pytest.mark.parametrize("scale", [1e308, 1e-309, 1.0])
def test_fp32_point_costs_need_a_normal_reciprocal_of_the_scale(scale):
    """FP32 storage forms the normalised point cost as raw * (1 / max): a
    finite maximum outside [2^-1022, 2^1022] (here L1 costs near 1e308 and
    near 1e-309, a subnormal) is rejected there; FP64 storage divides and
    accepts it."""
    g = np.random.default_rng(5)
    x, y = g.random((40, 1)) * scale, g.random((30, 1)) * scale
    prob = P.PointCloudProblem(x, y, np.ones(40) / 40, np.ones(30) / 30, "l1", normalize=True)
    mx = np.abs(y[:, 0][None, :] - x[:, 0][:, None]).max()
    assert np.isfinite(mx) and mx > 0
    out_of_range = not (2.0 ** -1022 <= mx <= 2.0 ** 1022)
    assert out_of_range == (scale != 1.0)
    d32 = _lib.DeviceSolver(40, 30, storage=_lib.STORAGE_F32)
    if out_of_range:
        with pytest.raises(_lib.PdotError, match=r"FP32 storage needs the normalising maximum"):
            d32.upload_points(prob)
    else:
        d32.upload_points(prob)
    d32.close()
    d64 = _lib.DeviceSolver(40, 30, storage=_lib.STORAGE_F64)
    d64.upload_points(prob)
    d64.close()


def test_round_matrix_rejects_bad_shapes():
    prob = instances.uniform_sqeuclid(10, 8, seed=1)
    with pytest.raises(ValueError):
        P.rounding(prob, np.zeros((8, 10)))


def test_state_and_defer_arguments_are_validated():
    prob = instances.uniform_sqeuclid(30, 20, seed=2)
    dev = _lib.DeviceSolver(30, 20)
    dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    dev.set_start()
    st = dev.run_steps(2)  # the second pass leaves a lagging contribution
    assert st.avg_pending == 1
    for field, value, msg in (("cur_slot", 4, "slot index"), ("spare1", -1, "slot index"),
                              ("avg_pending", 2, "0 or 1"), ("defer_ok", -1, "0 or 1")):
        bad = _lib.SolverState.from_buffer_copy(st)
        setattr(bad, field, value)
        with pytest.raises(_lib.PdotError, match=msg):
            dev.set_state(bad)
    bad = _lib.SolverState.from_buffer_copy(st)
    bad.avg_slot = bad.cur_slot
    with pytest.raises(_lib.PdotError, match="lagging contribution"):
        dev.set_state(bad)
    with pytest.raises(_lib.PdotError, match="0 or 1"):
        dev._check(dev.lib.pdot_set_defer(dev.h, 2), "pdot_set_defer")
    dev.set_state(st)
    assert dev.get_state().avg_pending == 1
    dev.close()

    probs = [instances.uniform_sqeuclid(n, m, seed=5 + n) for n, m in ((12, 9), (7, 15))]
    b = _lib.BatchSolver([p.num_row for p in probs], [p.num_col for p in probs], storage=_lib.STORAGE_F64)
    for k, p in enumerate(probs):
        b.upload(k, p.cost_matrix, p.source_distribution, p.target_distribution)
    b.set_start(_lib.POLICY_ADAPTIVE, 2.0, 1.0, 1.0, 1.0)
    states = b.run_steps([2, 3])
    assert [s.avg_pending for s in states] == [1, 0]
    for k, field, value, msg in ((1, "avg_slot", 9, "slot index"), (1, "avg_pending", 3, "0 or 1"),
                                 (0, "avg_count", 0, "lagging contribution")):
        bad = (_lib.SolverState * 2).from_buffer_copy(states)
        setattr(bad[k], field, value)
        with pytest.raises(_lib.PdotError, match=msg):
            b._check(b.lib.pdot_batch_set_states(b.h, bad), "pdot_batch_set_states")
    with pytest.raises(_lib.PdotError, match="0 or 1"):
        b._check(b.lib.pdot_batch_set_defer(b.h, 5), "pdot_batch_set_defer")
    assert [s.avg_pending for s in b.get_states()] == [1, 0]
    b.close()
