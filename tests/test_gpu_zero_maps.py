"""Zero maps (pdot_set_sparse): every iterate slot records, per 16 bytes of a
column, whether those entries are +0.0, and a pass neither reads nor writes
the entries recorded so.  The arithmetic is the same with the maps on or
off, so everything must agree bit for bit:

* whole solves of every form (dense and point costs, FP64 and FP32 storage,
  the per-kind kernels, deferral on and off), with restarts to the average
  (which writes out a lagging contribution over whole slots);
* iterates downloaded in the middle of a solve, across a restart to the
  average, a change of launch geometry (the maps are rebuilt) and the maps
  switched off and on again;
* the share of entries the maps skip equals the share of 16-byte units of
  the downloaded iterate that are +0.0 (the maps record exactly the zeros
  of a point written with them on);
* the batched passes (pdot_batch_set_sparse): whole batch solves, and
  iterates downloaded across restarts, a change of column tiling and a
  toggle of the maps."""
import numpy as np
import pytest

import pdot_amd as P
from pdot_amd import _lib, instances
from pdot_amd.batch import optimize_batch
from test_gpu_kernel_choices import _same_solves

pytestmark = pytest.mark.gpu

CASES = [("dense", _lib.STORAGE_F64, lambda: instances.gaussian_shift(700, 650, seed=31), 1e-6),
         ("dense", _lib.STORAGE_F32, lambda: instances.gaussian_shift(517, 260, seed=32), 1e-5),
         ("points", _lib.STORAGE_F32, lambda: instances.l1_mixture_3d_points(333, 210, seed=33), 1e-4),
         ("points", _lib.STORAGE_F64, lambda: instances.gaussian_shift_points(301, 299, seed=34), 1e-5),
         ("dense", _lib.STORAGE_F64, lambda: instances.gaussian_shift(4500, 4400, seed=24), 1e-4)]


def _solve(prob, storage, points, params, sparse, defer=True):
    n, m = len(prob.source_distribution), len(prob.target_distribution)
    dev = _lib.DeviceSolver(n, m, storage=storage)
    if points:
        dev.upload_points(prob)
    else:
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    dev.set_sparse(sparse)
    dev.set_defer(defer)
    out = P.optimize(params, prob, storage=storage, solver=dev)
    dev.close()
    return out


@pytest.mark.parametrize("form,storage,make,tol", CASES)
def test_zero_maps_agree_bitwise(form, storage, make, tol):
    prob = make()
    params = P.default_parameters(tol)
    _same_solves([_solve(prob, storage, form == "points", params, sparse) for sparse in (True, False)])


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_zero_maps_agree_bitwise_without_deferral(storage):
    prob = instances.gaussian_shift(640, 600, seed=35)
    params = P.default_parameters(1e-5)
    _same_solves([_solve(prob, storage, False, params, sparse, defer=False) for sparse in (True, False)])


def _pair(storage, prob):
    devs = []
    for sparse in (True, False):
        dev = _lib.DeviceSolver(prob.num_row, prob.num_col, storage=storage)
        dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
        dev.set_sparse(sparse)
        dev.set_start(policy=_lib.POLICY_ADAPTIVE, step_size=1.0, primal_weight=1.0)
        devs.append(dev)
    return devs


def _same_points(devs, avg=True):
    for which in ((_lib.POINT_CURRENT, _lib.POINT_AVERAGE) if avg else (_lib.POINT_CURRENT,)):
        (Xa, ua, va), (Xb, ub, vb) = (d.download(which) for d in devs)
        assert np.array_equal(Xa, Xb) and np.array_equal(ua, ub) and np.array_equal(va, vb)
    return devs[0].download(_lib.POINT_CURRENT)[0]


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_iterates_agree_across_restarts_geometry_changes_and_toggles(storage):
    prob = instances.gaussian_shift(900, 700, seed=36)
    devs = _pair(storage, prob)
    for d in devs:
        d.run_steps(37)  # odd: the average has a lagging contribution
    _same_points(devs)
    for d in devs:
        d.restart(to_average=True, primal_weight=0.7)
        d.run_steps(20)
    _same_points(devs)
    for d in devs:
        d.set_launch(col_blocks_hint=3)  # the maps are rebuilt for the new geometry
        d.run_steps(15)
    _same_points(devs)
    devs[0].set_sparse(False)
    devs[0].run_steps(4)
    devs[1].run_steps(4)
    devs[0].set_sparse(True)
    for d in devs:
        d.run_steps(11)
    X = _same_points(devs)
    assert np.array_equal(devs[0].round(_lib.POINT_AVERAGE), devs[1].round(_lib.POINT_AVERAGE))
    assert devs[1].zero_share() == 0.0  # maps off
    for d in devs:
        d.close()
    assert (X == 0).mean() > 0.5


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_zero_share_counts_the_zero_units(storage):
    prob = instances.gaussian_shift(1000, 800, seed=37)
    dev = _lib.DeviceSolver(1000, 800, storage=storage)
    dev.upload(prob.cost_matrix, prob.source_distribution, prob.target_distribution)
    dev.set_start(policy=_lib.POLICY_ADAPTIVE, step_size=1.0, primal_weight=1.0)
    assert dev.zero_share() == 1.0  # the start is all zeros
    dev.run_steps(300)
    share = dev.zero_share()
    X = dev.download(_lib.POINT_CURRENT)[0]
    vec = 2 if storage == _lib.STORAGE_F64 else 4
    n = X.shape[0]
    pad = np.zeros(((n + vec - 1) // vec * vec, X.shape[1]))
    pad[:n] = X
    units_zero = ~np.any(pad.reshape(-1, vec, X.shape[1]) != 0, axis=1)
    assert 0.3 < share < 1.0
    assert share == pytest.approx(units_zero.mean(), rel=0, abs=1e-12)
    dev.close()


def _batch_solve(probs, storage, params, sparse):
    b = _lib.BatchSolver([p.num_row for p in probs], [p.num_col for p in probs], storage=storage)
    for k, p in enumerate(probs):
        b.upload(k, p.cost_matrix, p.source_distribution, p.target_distribution)
    b.set_sparse(sparse)
    outs = optimize_batch(params, probs, storage=storage, batch=b)
    b.close()
    return outs


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_batch_zero_maps_agree_bitwise(storage):
    """The batched passes with zero maps (the default) against the same
    passes with the maps off, over uneven instances that finish at different
    times (the column tiling changes as instances leave)."""
    sizes = [(300, 257), (1, 40), (50, 1), (700, 650), (130, 1500), (64, 64)]
    probs = [instances.gaussian_shift(n, m, seed=40 + k) for k, (n, m) in enumerate(sizes)]
    params = P.default_parameters(1e-5)
    on, off = (_batch_solve(probs, storage, params, s) for s in (True, False))
    for a, b in zip(on, off):
        assert a.termination_reason == P.TerminationReason.TERMINATION_REASON_OPTIMAL
        assert b.iteration_count == a.iteration_count
        assert [s.restart_used for s in a.iteration_stats] == [s.restart_used for s in b.iteration_stats]
        for sa, sb in zip(a.iteration_stats, b.iteration_stats):
            assert sa.convergence_information[0].primal_objective == sb.convergence_information[0].primal_objective
        assert np.array_equal(a.primal_solution, b.primal_solution)
        assert np.array_equal(a.dual_source_solution, b.dual_source_solution)
        assert np.array_equal(a.dual_target_solution, b.dual_target_solution)
    assert max(o.iteration_count for o in on) > 4 * min(o.iteration_count for o in on)


@pytest.mark.parametrize("storage", [_lib.STORAGE_F64, _lib.STORAGE_F32])
def test_batch_iterates_agree_across_restarts_tilings_and_toggles(storage):
    """Batched iterates with the maps on and off, downloaded mid-solve: after
    a restart to the average with lagging contributions, a change of column
    tiling (the maps follow the fixed row tiles, not the columns) and the
    maps switched off (every slot expanded in place) and on again."""
    sizes = [(300, 257), (700, 650), (64, 900)]
    probs = [instances.gaussian_shift(n, m, seed=60 + k) for k, (n, m) in enumerate(sizes)]
    bs = []
    for sparse in (True, False):
        b = _lib.BatchSolver([n for n, _ in sizes], [m for _, m in sizes], storage=storage)
        for k, p in enumerate(probs):
            b.upload(k, p.cost_matrix, p.source_distribution, p.target_distribution)
        b.set_sparse(sparse)
        b.set_start(_lib.POLICY_ADAPTIVE, 2.0, 1.0, 1.0, [1.0, 0.5, 2.0])
        bs.append(b)

    def same(average=True):
        for k in range(len(sizes)):
            for which in ((_lib.POINT_CURRENT, _lib.POINT_AVERAGE) if average else (_lib.POINT_CURRENT,)):
                xs, ys = (b.download(k, which) for b in bs)
                for x, y in zip(xs, ys):
                    assert np.array_equal(x, y)
        return bs[0].download(1, _lib.POINT_CURRENT)[0]

    for b in bs:
        b.run_steps([41, 37, 30])
    same()
    for b in bs:
        b.restart([2, 2, 1], [0.7, 1.3, 1.0])
        b.run_steps([20, 21, 22])
    same()
    for b in bs:
        b.set_launch(32, 1)
        b.run_steps([15, 15, 15])
    same()
    bs[0].set_sparse(False)
    for b in bs:
        b.run_steps([5, 4, 3])
    bs[0].set_sparse(True)
    for b in bs:
        b.run_steps([11, 12, 13])
    X = same()
    for b in bs:
        b.close()
    assert (X == 0).mean() > 0.5
