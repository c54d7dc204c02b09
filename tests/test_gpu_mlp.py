"""GPU tests of the warm-start network kernels (pdot_mlp): the bf16 MFMA
matrix product, forward pass, loss gradient and Adam step against the numpy
restatement (oracle/mlp_oracle.py) with the same bf16 roundings."""
import numpy as np
import pytest

import mlp_oracle as M
from pdot_amd import learned

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    n = learned.WarmStartNet(seed=3)
    yield n
    n.close()


@pytest.mark.parametrize("M_,N_,K_,ta,tb", [(32, 32, 16, 0, 0), (37, 45, 21, 0, 0), (64, 128, 128, 0, 0),
                                            (128, 128, 300, 1, 0), (300, 128, 128, 0, 1), (5, 3, 7, 1, 1),
                                            (16, 128, 1000, 1, 0), (1000, 1, 128, 0, 0),
                                            # split over K (few output tiles, long K), as the weight gradients
                                            (128, 128, 8192, 1, 0), (64, 32, 5000, 0, 0), (128, 1, 8191, 1, 0)])
def test_mfma_product_matches_emulation(net, M_, N_, K_, ta, tb):
    rng = np.random.default_rng(M_ + N_ + K_)
    A = rng.standard_normal((K_, M_) if ta else (M_, K_)).astype(np.float32)
    B = rng.standard_normal((N_, K_) if tb else (K_, N_)).astype(np.float32)
    C = net.gemm_probe(A, B, bool(ta), bool(tb))
    ref = M.gemm(A, B, bool(ta), bool(tb))
    np.testing.assert_allclose(C, ref, rtol=2e-5, atol=2e-5 * np.abs(ref).max())


def test_mfma_lane_layout_identity(net):
    """An exact case: 0/1 operands are bf16-exact and all sums are small
    integers, so every output element must be bit-exact."""
    rng = np.random.default_rng(7)
    A = (rng.random((64, 48)) < 0.3).astype(np.float32)
    B = (rng.random((48, 96)) < 0.3).astype(np.float32)
    assert np.array_equal(net.gemm_probe(A, B), (A.astype(np.float64) @ B).astype(np.float32))


@pytest.mark.parametrize("ns,nt", [(23, 17), (300, 257), (1, 64), (700, 3)])
def test_forward_and_gradient_match_oracle(net, ns, nt):
    rng = np.random.default_rng(ns * 31 + nt)
    params = learned.init_params(seed=ns)
    params[np.arange(params.size) % 7 == 0] += 0.05  # non-zero biases too
    net.set_params(params)
    feats = rng.random((ns + nt, 4)).astype(np.float32)
    target = (0.3 * rng.standard_normal(ns + nt)).astype(np.float32)
    out = net.forward(feats, ns, nt)
    ref_out, _ = M.forward(params.astype(np.float64), feats, ns, nt, emulate=True)
    # The GPU forms pre-activations in FP32 and the emulation in FP64; where
    # the two straddle a bf16 rounding boundary, the next product's operand
    # differs by one bf16 ulp (2^-8 relative).  Expected flips are a few per
    # 10^5 activations, so almost all outputs agree to 1e-4 and the rest to a
    # few bf16 ulps of one summand.
    scale = np.abs(ref_out).max()
    err = np.abs(out - ref_out)
    assert np.mean(err > 1e-4 * scale) <= 0.01
    assert err.max() <= 2e-3 * scale
    loss, grad = net.loss_grad(feats, ns, nt, target, want_grad=True)
    rl, rg, _ = M.loss_grad(params.astype(np.float64), feats, ns, nt, target, emulate=True)
    assert loss == pytest.approx(rl, rel=1e-4)
    assert np.linalg.norm(grad - rg) / np.linalg.norm(rg) < 1e-3
    # every block, not just the norm
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        g, r = grad[o:o + k], rg[o:o + k]
        assert np.linalg.norm(g - r) <= 2e-3 * max(np.linalg.norm(r), 1e-6 * np.linalg.norm(rg)), name
        o += k


def test_adam_step_matches_numpy(net):
    rng = np.random.default_rng(9)
    params = learned.init_params(seed=9)
    net.set_params(params)
    feats = rng.random((50, 4)).astype(np.float32)
    target = rng.standard_normal(50).astype(np.float32)
    _, g = net.loss_grad(feats, 30, 20, target, want_grad=True)
    net.adam_step(1e-3)
    p1 = net.get_params()
    m = 0.1 * g.astype(np.float64)
    v = 0.001 * g.astype(np.float64) ** 2
    ref = params - 1e-3 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    np.testing.assert_allclose(p1, ref, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss(net):
    rng = np.random.default_rng(11)
    net.set_params(learned.init_params(seed=11))
    samples = []
    for k in range(8):
        x, y = rng.random((60, 2)), rng.random((50, 2))
        a = rng.dirichlet(np.ones(60))
        b = rng.dirichlet(np.ones(50))
        u, v = x.sum(axis=1) - 1.0, -(y[:, 0] ** 2)  # a smooth function of position to fit
        samples.append(learned.make_sample(x, y, a, b, u, v))
    before = learned.evaluate_loss(net, samples)
    learned.train(net, samples, epochs=30, lr=3e-3)
    after = learned.evaluate_loss(net, samples)
    assert after < 0.2 * before


@pytest.mark.parametrize("ns,nt,train_trunk", [(23, 17, False), (300, 257, True), (1, 64, True), (700, 3, False)])
def test_weight_head_matches_oracle(net, ns, nt, train_trunk):
    """The weight head's prediction, loss and gradient (FP32 head on the
    bf16-MFMA first layer) against the restatement with the same roundings."""
    rng = np.random.default_rng(ns * 13 + nt)
    params = learned.init_params(seed=ns + 1)
    params[np.arange(params.size) % 5 == 0] += 0.05
    net.set_params(params)
    feats = rng.random((ns + nt, 4)).astype(np.float32)
    lw = net.log_weight(feats, ns, nt)
    ref, _ = M.forward_weight(params.astype(np.float64), feats, ns, nt, emulate=True)
    assert lw == pytest.approx(ref, rel=1e-4, abs=1e-5)
    loss, pred, grad = net.weight_loss_grad(feats, ns, nt, -0.4, train_trunk, want_grad=True)
    rl, rg, rout = M.weight_loss_grad(params.astype(np.float64), feats, ns, nt, -0.4, train_trunk, emulate=True)
    assert pred == pytest.approx(rout, rel=1e-4, abs=1e-5)
    assert loss == pytest.approx(rl, rel=1e-3, abs=1e-6)
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        g, r = grad[o:o + k], rg[o:o + k]
        if not np.any(r):
            assert not np.any(g), name
        else:
            assert np.linalg.norm(g - r) <= 2e-3 * np.linalg.norm(r), name
        o += k


def test_dual_loss_leaves_weight_head_alone(net):
    """After a weight-loss gradient, the dual loss's gradient is zero on the
    head (no stale head gradient reaches Adam)."""
    rng = np.random.default_rng(5)
    net.set_params(learned.init_params(seed=5))
    feats = rng.random((90, 4)).astype(np.float32)
    net.weight_loss_grad(feats, 40, 50, 1.0, True)
    _, g = net.loss_grad(feats, 40, 50, rng.standard_normal(90).astype(np.float32), want_grad=True)
    assert not np.any(g[learned.NUM_DUAL_PARAMS:])
    assert np.any(g[:learned.NUM_DUAL_PARAMS])


def test_weight_head_training_fits_targets(net):
    """Adam on the weight loss (head only) fits log weights that depend on
    the point clouds' spread; the dual head's parameters do not move."""
    rng = np.random.default_rng(21)
    p0 = learned.init_params(seed=21)
    net.set_params(p0)
    data = []
    for k in range(12):
        spread = 0.2 + 0.8 * k / 11
        x, y = spread * rng.random((80, 2)), spread * rng.random((70, 2))
        a, b = rng.dirichlet(np.ones(80)), rng.dirichlet(np.ones(70))
        data.append((learned.features(x, y, a, b), 80, 70, float(np.log(0.05 + 0.5 * spread))))
    before = np.mean([net.weight_loss_grad(f, ns, nt, t)[0] for f, ns, nt, t in data])
    for ep in range(150):
        for f, ns, nt, t in data:
            net.weight_loss_grad(f, ns, nt, t, False)
            net.adam_step(3e-3)
    after = np.mean([net.weight_loss_grad(f, ns, nt, t)[0] for f, ns, nt, t in data])
    assert after < 0.1 * before
    p1 = net.get_params()
    assert np.array_equal(p1[:learned.NUM_DUAL_PARAMS], p0[:learned.NUM_DUAL_PARAMS])
