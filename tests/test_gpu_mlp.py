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
                                            (16, 128, 1000, 1, 0), (1000, 1, 128, 0, 0)])
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
