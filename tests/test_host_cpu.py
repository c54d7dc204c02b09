"""CPU tests of host-side logic that needs no GPU: the warm-start feature and
target construction, the closed-form first-evaluation KKT error of the zero
start used by the warm start's acceptance test (pinned against the oracle's first
evaluation), the point-cloud problem's dense matrix, and the batch tile-width
rule's inputs."""
import json
import os

import numpy as np
import pytest

import oracle as O
import pdot_amd as P
from pdot_amd import instances, learned

REPO_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_normalize_duals_keeps_the_dual_objective_and_balances():
    rng = np.random.default_rng(1)
    a, b = rng.dirichlet(np.ones(30)), rng.dirichlet(np.ones(20))
    u, v = rng.standard_normal(30), rng.standard_normal(20)
    un, vn = learned.normalize_duals(u, v, a, b)
    assert a @ un + b @ vn == pytest.approx(a @ u + b @ v, rel=1e-14)
    assert a @ un == pytest.approx(b @ vn, abs=1e-14)
    # invariant under the shift it removes
    us, vs = learned.normalize_duals(u + 3.0, v - 3.0, a, b)
    np.testing.assert_allclose(us, un, atol=1e-13)
    np.testing.assert_allclose(vs, vn, atol=1e-13)


def test_features_layout():
    x, y = np.array([[0.1, 0.2], [0.3, 0.4]]), np.array([[0.5, 0.6], [0.7, 0.8], [0.9, 1.0]])
    a, b = np.array([0.25, 0.75]), np.array([0.2, 0.3, 0.5])
    f = learned.features(x, y, a, b)
    assert f.shape == (5, 4) and f.dtype == np.float32
    np.testing.assert_allclose(f[:2, :2], x)
    np.testing.assert_allclose(f[2:, :2], y)
    np.testing.assert_allclose(f[:2, 2], 2 * a)
    np.testing.assert_allclose(f[2:, 2], 3 * b)
    assert list(f[:, 3]) == [1, 1, -1, -1, -1]


@pytest.mark.parametrize("seed", [1, 2])
def test_zero_start_kkt_closed_form_matches_oracle(seed):
    prob = instances.uniform_sqeuclid(40, 33, seed=seed)
    r = O.optimize(P.default_parameters(1e-4, iteration_limit=1), prob, mode=O.FUSED)
    s = r["stats"][0]
    kkt = max(s["relative_l2_primal_residual"], s["relative_l2_dual_residual"], s["relative_optimality_gap"])
    assert learned.zero_start_relative_kkt(prob) == pytest.approx(kkt, rel=1e-14)


def test_point_cloud_dense_matches_generators():
    for make_pts, make_dense, args in ((instances.gaussian_shift_points, instances.gaussian_shift, (60, 45, 3)),
                                       (instances.l1_mixture_3d_points, instances.l1_mixture_3d, (50, 40, 2)),
                                       (instances.uniform_sqeuclid_points, instances.uniform_sqeuclid, (30, 20, 5))):
        d = make_pts(*args).dense()
        e = make_dense(*args)
        assert np.array_equal(d.cost_matrix, e.cost_matrix)
        assert np.array_equal(d.source_distribution, e.source_distribution)


def test_point_cloud_problem_validation():
    with pytest.raises(ValueError):
        P.PointCloudProblem(np.zeros((3, 4)), np.zeros((2, 4)), np.ones(3) / 3, np.ones(2) / 2)
    with pytest.raises(ValueError):
        P.PointCloudProblem(np.zeros((3, 2)), np.zeros((2, 2)), np.ones(3) / 3, np.ones(2) / 2, cost="cosine")
    with pytest.raises(ValueError):
        P.PointCloudProblem(np.zeros((3, 2)), np.zeros((2, 3)), np.ones(3) / 3, np.ones(2) / 2)


def test_calibrated_primal_weight_is_stored_with_the_network():
    """The committed weights carry the calibration of the initial primal
    weight and the weight head's selection, consistent with the committed
    training report: the constant is a grid value (not the reference's 1.0)
    chosen over at least 20,000 grid solves, and the head is used only if it
    needs at most SELECT_MARGIN times the constant's validation iterations."""
    import json
    from pdot_amd import train_warm_start as T
    meta = learned.load_meta()
    report = json.load(open(os.path.join(os.path.dirname(learned.DEFAULT_WEIGHTS), "warm_start_report.json")))
    w = learned.load_initial_primal_weight()
    grid = [float(g) for g in meta["weight_grid"]]
    assert w in grid and 1.0 not in grid
    assert w == report["calibrated_constant"]
    assert int(meta["weight_grid_solves"]) == report["grid_solves"] >= 20000
    selected = bool(int(meta["weight_head_selected"]))
    assert selected == report["head_selected"]
    head, const = float(meta["weight_val_iterations_head"]), float(meta["weight_val_iterations_constant"])
    assert (head, const) == (report["val_iterations_head"], report["val_iterations_constant"])
    assert selected == (head <= T.SELECT_MARGIN * const)


def test_point_third_evaluation_known_answer_matches_oracle():
    """The host known answer used by the large point-cloud GPU test agrees
    with the oracle's iteration-3 evaluation on a small problem."""
    from point_kat import third_evaluation
    g = np.random.default_rng(3)
    pts = P.PointCloudProblem(g.normal(size=(130, 2)), g.normal(size=(97, 2)) + [2.0, 0.0],
                              g.random(130) / 65, g.random(97) / 48.5, "sqeuclid", normalize=False)
    params = P.default_parameters(1e-4, iteration_limit=3)
    ref = O.optimize(params, pts.dense(), mode=O.FUSED)
    norm, obj = third_evaluation(pts, block=40)
    assert ref["stats"][2]["l2_primal_variable"] == pytest.approx(norm, rel=1e-13)
    assert ref["stats"][2]["primal_objective"] == pytest.approx(obj, rel=1e-13)


def test_bench_byte_accounting_matches_the_traffic_profiles():
    """The bench's byte accounting against the calibrated HBM traffic recorded
    in profiles/ at the same state of the 10k solve: the dense figure (4
    element streams on average with deferred averaging, 3 for point clouds)
    and the bytes the passes must move with zero maps at the recorded zero
    shares.  The traffic lies between the two: at least the needed bytes
    (partial sums and cache-line rounding on top, under 30%), far below dense."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO_ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    n = 10000
    assert bench.algorithmic_bytes_per_step(n, n, 8) == 4 * n * n * 8 + 80 * 2 * n
    assert bench.algorithmic_bytes_per_step(n, n, 8, matrix_streams=3) == 3 * n * n * 8 + 80 * 2 * n
    with open(os.path.join(REPO_ROOT, "profiles", "hbm_traffic_step_kernel_gaussian_shift_10k.md")) as f:
        rec = json.loads(f.read().split("```json", 1)[1].split("```", 1)[0])
    for form, elem, streams in (("f64", 8, 4), ("f32", 4, 4), ("f64_points", 8, 3), ("f32_points", 4, 3)):
        r = rec[form]
        traffic = bench.measured_traffic(n, n, form)
        assert traffic == r["traffic_bytes"]
        dense = bench.algorithmic_bytes_per_step(n, n, elem, matrix_streams=streams)
        needed = bench.sparse_bytes_per_step(r["launch"], n, n, elem, streams == 4, r["zero_share_current"],
                                             r["zero_share_average"])
        # no zeros: the dense figure plus the maps
        nozero = bench.sparse_bytes_per_step(r["launch"], n, n, elem, streams == 4, 0.0, 0.0)
        assert dense < nozero < 1.01 * dense, form
        assert 1.0 <= traffic / needed < 1.3, form
        assert traffic < 0.5 * dense, form
    assert bench.measured_traffic(5000, 5000, "f64") is None
    bt = bench.measured_batch_traffic()
    assert 8.8e9 < bt < 9.2e9


def test_best_log_weight_parabola_and_edges():
    """The weight target: the grid minimum of the iteration count, refined by
    the parabola through log(iterations) around it (vertex of an exact
    parabola recovered), kept at the minimum on the grid's edge or next to a
    capped run, NaN when no run reached the tolerance."""
    import math
    from pdot_amd import train_warm_start as T
    grid = (0.2, 0.1, 0.4, 0.05, 0.8)  # run order, not sorted
    lw = np.log(np.array(grid))
    vertex = math.log(0.15)
    it = np.exp(2.0 * (lw - vertex) ** 2 + 7.0)
    assert T.best_log_weight(grid, it) == pytest.approx(vertex, abs=1e-9)
    # minimum at the edge of the grid
    it2 = np.array([500, 400, 600, 300, 700])
    assert T.best_log_weight(grid, it2) == pytest.approx(math.log(0.05))
    # a capped neighbour (-1) keeps the grid minimum
    it3 = np.array([300, 200, 500, -1, 700])
    assert T.best_log_weight(grid, it3) == pytest.approx(math.log(0.1))
    assert math.isnan(T.best_log_weight(grid, np.array([-1, -1, -1, -1, -1])))


def test_calibrated_constant_minimises_total_iterations():
    from pdot_amd import train_warm_start as T
    data = {"grid": np.array([0.2, 0.1, 0.4]),
            "iterations": np.array([[100, 300, 90], [200, 150, -1], [400, 100, 500]])}
    # capped runs count as 3x the instance's best: totals 700, 550, 90 + 450 + 500 = 1040
    assert T.calibrated_constant(data) == 0.1
