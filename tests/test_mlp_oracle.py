"""Pins the numpy restatement of the warm-start network (oracle/mlp_oracle.py)
before it checks the GPU kernels: bf16 rounding known answers, and the
backward pass against central finite differences of the plain FP64
network (a dropped term, wrong sign or transposed operand fails them)."""
import numpy as np
import pytest

import mlp_oracle as M


def test_bf16_rounding_known_answers():
    one = 1.0
    assert M.bf16(one)[()] == 1.0
    assert M.bf16(1.0 + 2.0 ** -8)[()] == 1.0                   # tie -> even
    assert M.bf16(1.0 + 3 * 2.0 ** -9)[()] == 1.0 + 2.0 ** -7    # above the tie
    assert M.bf16(1.0 + 2.0 ** -7 + 2.0 ** -8)[()] == 1.0 + 2.0 ** -6  # tie -> even (up)
    assert M.bf16(-2.5)[()] == -2.5
    x = np.random.default_rng(0).standard_normal(1000)
    r = M.bf16(x)
    assert np.all(np.abs(r - x) <= np.abs(x) * 2.0 ** -8 * 1.0001)


def _problem(seed, ns=23, nt=17):
    rng = np.random.default_rng(seed)
    feats = rng.standard_normal((ns + nt, M.F))
    params = rng.standard_normal(M.NUM_PARAMS) * 0.2
    target = rng.standard_normal(ns + nt)
    return params, feats, target, ns, nt


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_gradient_matches_finite_differences(seed):
    params, feats, target, ns, nt = _problem(seed)
    loss, grad, _ = M.loss_grad(params, feats, ns, nt, target, emulate=False)
    rng = np.random.default_rng(100 + seed)
    # probe coordinates in every parameter block
    offsets, o = [], 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        offsets += list(o + rng.choice(k, size=min(k, 6), replace=False))
        o += k
    h = 1e-6
    for i in offsets:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (M.loss_grad(p1, feats, ns, nt, target, emulate=False)[0]
              - M.loss_grad(p2, feats, ns, nt, target, emulate=False)[0]) / (2 * h)
        assert grad[i] == pytest.approx(fd, rel=1e-5, abs=1e-8), i


def test_emulated_numerics_stay_close_to_fp64():
    params, feats, target, ns, nt = _problem(3)
    l64, g64, _ = M.loss_grad(params, feats, ns, nt, target, emulate=False)
    lbf, gbf, _ = M.loss_grad(params, feats, ns, nt, target, emulate=True)
    assert lbf == pytest.approx(l64, rel=2e-2)
    assert np.linalg.norm(gbf - g64) / np.linalg.norm(g64) < 5e-2


@pytest.mark.parametrize("seed,train_trunk", [(0, False), (1, True), (2, True)])
def test_weight_head_gradient_matches_finite_differences(seed, train_trunk):
    """The weight head's loss (log_weight - target)^2: its restated gradient
    against central differences of the plain FP64 network, for the head's
    blocks and, when the trunk is trained, W1 and b1; every other block is
    zero."""
    params, feats, _, ns, nt = _problem(seed)
    target = 0.7
    loss, grad, out = M.weight_loss_grad(params, feats, ns, nt, target, train_trunk, emulate=False)
    assert loss == pytest.approx((out - target) ** 2)
    rng = np.random.default_rng(200 + seed)
    h = 1e-6
    o = 0
    for name, shape in M.SHAPES:
        k = int(np.prod(shape))
        live = name in M.HEAD or (train_trunk and name in ("W1", "b1"))
        if not live:
            assert not np.any(grad[o:o + k]), name
        else:
            for i in o + rng.choice(k, size=min(k, 6), replace=False):
                p1, p2 = params.copy(), params.copy()
                p1[i] += h
                p2[i] -= h
                fd = (M.weight_loss_grad(p1, feats, ns, nt, target, train_trunk, emulate=False)[0]
                      - M.weight_loss_grad(p2, feats, ns, nt, target, train_trunk, emulate=False)[0]) / (2 * h)
                assert grad[i] == pytest.approx(fd, rel=1e-5, abs=1e-8), (name, i)
        o += k


def test_weight_head_size_feature():
    assert M.size_feature(4096, 4096) == 1.0
    assert M.size_feature(1024, 2048) == pytest.approx(21 / 24)
