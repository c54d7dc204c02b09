"""numpy restatement of the warm-start network of libpdot_hip.so
(pdot_mlp.hip): forward pass, loss and gradient.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the GPU
kernels; product code never imports it.

emulate=True reproduces the GPU numerics: every matrix-product operand is
rounded FP64 -> FP32 -> bf16 (round to nearest even) and the product is then
accumulated in FP64 (the GPU accumulates in FP32); everything else is FP64.
emulate=False is the plain FP64 network, used to pin this restatement's
backward pass against central finite differences.
"""
from __future__ import annotations

import numpy as np

W = 128
F = 4
H = 32
SHAPES = [("W1", (F, W)), ("b1", (W,)), ("W2a", (W, W)), ("W2b", (W, W)), ("W2c", (W, W)), ("b2", (W,)),
          ("W3", (W, W)), ("b3", (W,)), ("w4", (W,)), ("b4", (1,)),
          ("Wh1a", (W, H)), ("Wh1b", (W, H)), ("wh1s", (H,)), ("bh1", (H,)), ("wh2", (H,)), ("bh2", (1,))]
HEAD = ("Wh1a", "Wh1b", "wh1s", "bh1", "wh2", "bh2")
NUM_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x) -> np.ndarray:
    """Round to bf16 (via FP32, nearest even), returned as float64."""
    x32 = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def pack(d) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def _mm(A, B, emulate):
    return bf16(A) @ bf16(B) if emulate else A @ B


def forward(params, feats, ns: int, nt: int, emulate: bool = True):
    P = unpack(params)
    X0 = np.asarray(feats, dtype=np.float64).reshape(ns + nt, F)
    Z1 = _mm(X0, P["W1"], emulate) + P["b1"]
    H1 = np.maximum(Z1, 0.0)
    cs = H1[:ns].mean(axis=0) if ns else np.zeros(W)
    ct = H1[ns:].mean(axis=0) if nt else np.zeros(W)
    g = (cs @ P["W2b"] + ct @ P["W2c"]) + P["b2"]
    Z2 = _mm(H1, P["W2a"], emulate) + g
    H2 = np.maximum(Z2, 0.0)
    Z3 = _mm(H2, P["W3"], emulate) + P["b3"]
    H3 = np.maximum(Z3, 0.0)
    out = _mm(H3, P["w4"][:, None], emulate)[:, 0] + P["b4"][0]
    cache = dict(P=P, X0=X0, Z1=Z1, H1=H1, cs=cs, ct=ct, Z2=Z2, H2=H2, Z3=Z3, H3=H3, ns=ns, nt=nt)
    return out, cache


def loss_grad(params, feats, ns: int, nt: int, target, emulate: bool = True):
    out, c = forward(params, feats, ns, nt, emulate)
    P = c["P"]
    N = ns + nt
    t = np.asarray(target, dtype=np.float64)
    d = out - t
    loss = float(np.mean(d * d))
    dout = 2.0 * d / N
    g = {}
    g["w4"] = _mm(c["H3"].T, dout[:, None], emulate)[:, 0]
    g["b4"] = np.array([dout.sum()])
    dZ3 = _mm(dout[:, None], P["w4"][None, :], emulate) * (c["Z3"] > 0)
    g["W3"] = _mm(c["H2"].T, dZ3, emulate)
    g["b3"] = dZ3.sum(axis=0)
    dZ2 = _mm(dZ3, P["W3"].T, emulate) * (c["Z2"] > 0)
    g["W2a"] = _mm(c["H1"].T, dZ2, emulate)
    dg = dZ2.sum(axis=0)
    g["b2"] = dg
    g["W2b"] = np.outer(c["cs"], dg)
    g["W2c"] = np.outer(c["ct"], dg)
    dcs, dct = P["W2b"] @ dg, P["W2c"] @ dg
    dH1 = _mm(dZ2, P["W2a"].T, emulate)
    if ns:
        dH1[:ns] += dcs / ns
    if nt:
        dH1[ns:] += dct / nt
    dZ1 = dH1 * (c["Z1"] > 0)
    g["W1"] = _mm(c["X0"].T, dZ1, emulate)
    g["b1"] = dZ1.sum(axis=0)
    for name in HEAD:  # the dual loss does not reach the weight head
        g[name] = np.zeros(dict(SHAPES)[name])
    return loss, pack(g), out


def size_feature(ns: int, nt: int) -> float:
    return (np.log2(max(ns, 1)) + np.log2(max(nt, 1))) / 24.0


def forward_weight(params, feats, ns: int, nt: int, emulate: bool = True):
    """The weight head: log of the initial primal weight.  The head itself is
    plain arithmetic (no bf16 operands); its inputs come from layer 1."""
    P = unpack(params)
    X0 = np.asarray(feats, dtype=np.float64).reshape(ns + nt, F)
    Z1 = _mm(X0, P["W1"], emulate) + P["b1"]
    H1 = np.maximum(Z1, 0.0)
    cs, ct = H1[:ns].mean(axis=0), H1[ns:].mean(axis=0)
    s = size_feature(ns, nt)
    zh = ((cs @ P["Wh1a"] + ct @ P["Wh1b"]) + s * P["wh1s"]) + P["bh1"]
    hh = np.maximum(zh, 0.0)
    out = float(hh @ P["wh2"] + P["bh2"][0])
    return out, dict(P=P, X0=X0, Z1=Z1, cs=cs, ct=ct, s=s, zh=zh, hh=hh)


def weight_loss_grad(params, feats, ns: int, nt: int, target: float, train_trunk: bool = False,
                     emulate: bool = True):
    """(log_weight - target)^2 and its gradient (zero outside the head, and
    outside the head and layer 1 when train_trunk)."""
    out, c = forward_weight(params, feats, ns, nt, emulate)
    P = c["P"]
    dl = 2.0 * (out - target)
    g = {name: np.zeros(shape) for name, shape in SHAPES}
    g["wh2"] = dl * c["hh"]
    g["bh2"] = np.array([dl])
    dz = dl * P["wh2"] * (c["zh"] > 0)
    g["bh1"] = dz
    g["wh1s"] = c["s"] * dz
    g["Wh1a"] = np.outer(c["cs"], dz)
    g["Wh1b"] = np.outer(c["ct"], dz)
    if train_trunk:
        dcs, dct = P["Wh1a"] @ dz, P["Wh1b"] @ dz
        dH1 = np.zeros((ns + nt, W))
        dH1[:ns] += dcs / ns
        dH1[ns:] += dct / nt
        dZ1 = dH1 * (c["Z1"] > 0)
        g["W1"] = _mm(c["X0"].T, dZ1, emulate)
        g["b1"] = dZ1.sum(axis=0)
    return (out - target) ** 2, pack(g), out


def gemm(A, B, trans_a=False, trans_b=False):
    """The MFMA product's numerics: bf16 operands, FP64 accumulation."""
    A = np.asarray(A, dtype=np.float64)
    B = np.asarray(B, dtype=np.float64)
    return bf16(A.T if trans_a else A) @ bf16(B.T if trans_b else B)
