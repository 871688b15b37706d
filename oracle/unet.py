"""NumPy FP64 restatement of the learned p' preconditioner (oracle, test-only).

The network is the 4-level U-Net of 3x3 convolutions (16 -> 32 -> 64 -> 128
channels, BF16 weights and activations, FP32 accumulation on the GPU) that
maps (p' residual, aP field, identity-row mask) to a p' correction.  The
reference has no such network; this file restates the product's definition
so that its HIP forward/backward kernels can be checked.  It is pinned by
torch CPU autograd (FP64, ``torch.nn.functional.conv2d``) and by central
finite differences in ``tests/test_oracle_unet.py``.

Architecture (channels-last tensors (B, H, W, C); H x W is the interior
nx x ny rounded up to multiples of 8, pixel (a, b) is cell (a + 1, b + 1)):

    X0  (3)                      features
    e0 = relu(conv(X0))          16, full resolution
    e1 = relu(conv(pool(e0)))    32, 1/2
    e2 = relu(conv(pool(e1)))    64, 1/4
    e3 = relu(conv(pool(e2)))   128, 1/8
    d2 = relu(conv(up(e3) | e2)) 64, 1/4
    d1 = relu(conv(up(d2) | e1)) 32, 1/2
    d0 = relu(conv(up(d1) | e0)) 16, full
    net = w_out . d0 + b_out     1 (1x1 convolution)

pool = 2x2 average, up = 2x nearest, ``|`` = channel concatenation (upsampled
channels first).  Features: r / rs, aP / abar on free cells, 1 on identity
rows (all four links zero), zero outside the interior; rs = ||r|| / sqrt(n_free),
abar = mean aP over free interior cells.  Output z = (rs / abar) net on free
interior cells, 0 elsewhere.

BF16 rounding points mirror the kernels when ``bf16=True``: features, conv
and 1x1 weights, every activation written by a layer (after ReLU), pooled
tensors, and the gradient tensors fed to the backward GEMMs.  Biases, the
gradient sums and the loss stay FP32/FP64.

Parameter layout (one flat FP32 vector): for each layer in the order
e0 e1 e2 e3 d2 d1 d0, W[cout][3][3][cin] then b[cout]; then w_out[16], b_out.
"""
from __future__ import annotations

import math

import numpy as np

from .krylov import stencil

LAYERS = (("e0", 3, 16), ("e1", 16, 32), ("e2", 32, 64), ("e3", 64, 128),
          ("d2", 192, 64), ("d1", 96, 32), ("d0", 48, 16))
C_OUT_IN = 16


def param_slices():
    """name -> (offset, shape) of every parameter in the flat vector."""
    out, off = {}, 0
    for name, cin, cout in LAYERS:
        out[name + ".w"] = (off, (cout, 3, 3, cin))
        off += cout * 9 * cin
        out[name + ".b"] = (off, (cout,))
        off += cout
    out["out.w"] = (off, (C_OUT_IN,))
    off += C_OUT_IN
    out["out.b"] = (off, (1,))
    off += 1
    return out, off


def param_count() -> int:
    return param_slices()[1]


def unpack(params):
    sl, _ = param_slices()
    return {k: np.asarray(params[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for k, (o, s) in sl.items()}


def pack(d):
    sl, n = param_slices()
    out = np.zeros(n)
    for k, (o, s) in sl.items():
        out[o:o + int(np.prod(s))] = np.asarray(d[k]).ravel()
    return out


def init_params(seed=0):
    """He-normal convolution weights, zero biases, N(0, 1/16) output weights."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, cin, cout in LAYERS:
        d[name + ".w"] = rng.standard_normal((cout, 3, 3, cin)) * math.sqrt(2.0 / (9 * cin))
        d[name + ".b"] = np.zeros(cout)
    d["out.w"] = rng.standard_normal(C_OUT_IN) * 0.25
    d["out.b"] = np.zeros(1)
    return pack(d).astype(np.float32)


def bf16(a):
    """Round to bfloat16 (round to nearest even) via FP32, returned as FP64."""
    f = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(f), f, out).astype(np.float64)


def _id(a):
    return np.asarray(a, dtype=np.float64)


def grid_dims(nx, ny):
    return -(-nx // 8) * 8, -(-ny // 8) * 8


def features(coef, r, bf=True):
    """(X0 (H, W, 3), rs, abar, free (nx, ny)) for one padded system."""
    NX, NY = r.shape
    nx, ny = NX - 2, NY - 2
    H, W = grid_dims(nx, ny)
    free_all = (coef[1] != 0) | (coef[2] != 0) | (coef[3] != 0) | (coef[4] != 0)
    free = free_all[1:-1, 1:-1]
    nfree = int(free.sum())
    ap = coef[0][1:-1, 1:-1]
    abar = float(ap[free].sum()) / nfree if nfree else 1.0
    rr = float(np.dot(r.ravel(), r.ravel()))
    rs = math.sqrt(rr / nfree) if nfree else 0.0
    X = np.zeros((H, W, 3))
    if rs > 0:
        X[:nx, :ny, 0] = (np.asarray(r[1:-1, 1:-1], np.float64) / rs).astype(np.float32)
    X[:nx, :ny, 1] = np.where(free, (ap / abar).astype(np.float32), 0.0)
    X[:nx, :ny, 2] = np.where(free, 0.0, 1.0)
    return (bf16(X) if bf else X), rs, abar, free


def conv3x3(x, w, b):
    """y[h, w, co] = b[co] + sum_{ty, tx, ci} x[h + ty - 1, w + tx - 1, ci] W[co, ty, tx, ci]."""
    B, H, Wd, C = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, C))
    xp[:, 1:-1, 1:-1] = x
    y = np.zeros((B, H, Wd, w.shape[0]))
    for ty in range(3):
        for tx in range(3):
            y += xp[:, ty:ty + H, tx:tx + Wd, :] @ w[:, ty, tx, :].T
    return y + b


def conv3x3_bwd_data(dy, w):
    B, H, Wd, _ = dy.shape
    dxp = np.zeros((B, H + 2, Wd + 2, w.shape[3]))
    for ty in range(3):
        for tx in range(3):
            dxp[:, ty:ty + H, tx:tx + Wd, :] += dy @ w[:, ty, tx, :]
    return dxp[:, 1:-1, 1:-1]


def conv3x3_wgrad(dy, x):
    B, H, Wd, C = x.shape
    xp = np.zeros((B, H + 2, Wd + 2, C))
    xp[:, 1:-1, 1:-1] = x
    dw = np.zeros((dy.shape[3], 3, 3, C))
    for ty in range(3):
        for tx in range(3):
            dw[:, ty, tx, :] = np.tensordot(dy, xp[:, ty:ty + H, tx:tx + Wd, :], axes=([0, 1, 2], [0, 1, 2]))
    return dw, dy.sum(axis=(0, 1, 2))


def pool(x):
    return 0.25 * (((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + x[:, 1::2, 0::2]) + x[:, 1::2, 1::2])


def unpool(d):
    """Adjoint of ``pool``."""
    return 0.25 * np.repeat(np.repeat(d, 2, axis=1), 2, axis=2)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def sum2x2(d):
    """Adjoint of ``up``."""
    return ((d[:, 0::2, 0::2] + d[:, 0::2, 1::2]) + d[:, 1::2, 0::2]) + d[:, 1::2, 1::2]


def forward(params, X0, bf=True):
    """Network output (B, H, W) and the saved activations."""
    R = bf16 if bf else _id
    P = unpack(params)
    w = {k: (R(v) if k.endswith(".w") else v) for k, v in P.items()}
    relu = lambda a: np.maximum(a, 0.0)
    c = {"X0": R(X0)}
    c["e0"] = R(relu(conv3x3(c["X0"], w["e0.w"], w["e0.b"])))
    c["p1"] = R(pool(c["e0"]))
    c["e1"] = R(relu(conv3x3(c["p1"], w["e1.w"], w["e1.b"])))
    c["p2"] = R(pool(c["e1"]))
    c["e2"] = R(relu(conv3x3(c["p2"], w["e2.w"], w["e2.b"])))
    c["p3"] = R(pool(c["e2"]))
    c["e3"] = R(relu(conv3x3(c["p3"], w["e3.w"], w["e3.b"])))
    c["x_d2"] = np.concatenate([up(c["e3"]), c["e2"]], axis=3)
    c["d2"] = R(relu(conv3x3(c["x_d2"], w["d2.w"], w["d2.b"])))
    c["x_d1"] = np.concatenate([up(c["d2"]), c["e1"]], axis=3)
    c["d1"] = R(relu(conv3x3(c["x_d1"], w["d1.w"], w["d1.b"])))
    c["x_d0"] = np.concatenate([up(c["d1"]), c["e0"]], axis=3)
    c["d0"] = R(relu(conv3x3(c["x_d0"], w["d0.w"], w["d0.b"])))
    net = c["d0"] @ w["out.w"] + w["out.b"][0]
    c["w"] = w
    return net, c


def backward(dnet, c, bf=True):
    """Gradient of sum(dnet * net) with respect to the flat parameters."""
    R = bf16 if bf else _id
    w = c["w"]
    g = {}
    g["out.w"] = np.tensordot(dnet, c["d0"], axes=([0, 1, 2], [0, 1, 2]))
    g["out.b"] = np.array([dnet.sum()])
    dy = R(dnet[..., None] * w["out.w"] * (c["d0"] > 0))
    # decoder: conv, then split the concatenated input gradient
    for name, x_key, coarse, skip_grad in (("d0", "x_d0", "d1", "s_e0"), ("d1", "x_d1", "d2", "s_e1"),
                                           ("d2", "x_d2", "e3", "s_e2")):
        g[name + ".w"], g[name + ".b"] = conv3x3_wgrad(dy, c[x_key])
        dx = conv3x3_bwd_data(dy, w[name + ".w"])
        cu = c[coarse].shape[3]
        g[skip_grad] = dx[..., cu:]
        dy = R(sum2x2(dx[..., :cu]) * (c[coarse] > 0))
    # encoder: conv on the pooled input, add the skip gradient, unpool
    for name, x_key, prev, skip_grad in (("e3", "p3", "e2", "s_e2"), ("e2", "p2", "e1", "s_e1"),
                                         ("e1", "p1", "e0", "s_e0")):
        g[name + ".w"], g[name + ".b"] = conv3x3_wgrad(dy, c[x_key])
        dp = conv3x3_bwd_data(dy, w[name + ".w"])
        dy = R((g.pop(skip_grad) + unpool(dp)) * (c[prev] > 0))
    g["e0.w"], g["e0.b"] = conv3x3_wgrad(dy, c["X0"])
    return pack(g)


def apply(params, coef, r, bf=True):
    """z = M(r) for one padded system (FP64 array of r's shape)."""
    X0, rs, abar, free = features(coef, r, bf)
    z = np.zeros(r.shape)
    if rs == 0.0:
        return z
    net, _ = forward(params, X0[None], bf)
    nx, ny = free.shape
    z[1:-1, 1:-1] = np.where(free, (rs / abar) * net[0, :nx, :ny], 0.0)
    return z


def stencil_t(coef, y):
    """A^T y for the 5-point operator of ``krylov.stencil``."""
    yp = np.zeros((y.shape[0] + 2, y.shape[1] + 2))
    yp[1:-1, 1:-1] = y
    c1 = np.zeros_like(yp); c1[1:-1, 1:-1] = coef[1]
    c2 = np.zeros_like(yp); c2[1:-1, 1:-1] = coef[2]
    c3 = np.zeros_like(yp); c3[1:-1, 1:-1] = coef[3]
    c4 = np.zeros_like(yp); c4[1:-1, 1:-1] = coef[4]
    return (coef[0] * y + (c1 * yp)[2:, 1:-1] + (c2 * yp)[:-2, 1:-1]
            + (c3 * yp)[1:-1, :-2] + (c4 * yp)[1:-1, 2:])


def loss_and_grad(params, coefs, rs_batch, bf=True, corrs=None):
    """Mean over the batch of the relative loss of z = M(r) and its gradient.

    Without ``corrs``: the residual loss ||r - A z||^2 / ||r||^2.
    With ``corrs`` (c = A^-1 r, the converged corrections): the A-norm error
    loss e^T A e / c^T r with e = c - z (= ||e||_A^2 / ||c||_A^2 for SPD A),
    whose gradient is -(A + A^T) e / c^T r.
    coefs: (B, 5, NX, NY); rs_batch, corrs: (B, NX, NY).  Returns (loss, grad, losses).
    """
    if corrs is not None:
        return _anorm_loss_and_grad(params, coefs, rs_batch, corrs, bf)
    Bn = len(rs_batch)
    feats = [features(coefs[k], rs_batch[k], bf) for k in range(Bn)]
    X0 = np.stack([f[0] for f in feats])
    net, cache = forward(params, X0, bf)
    dnet = np.zeros_like(net)
    losses = []
    for k in range(Bn):
        _, rs, abar, free = feats[k]
        r = np.asarray(rs_batch[k], np.float64)
        nx, ny = free.shape
        z = np.zeros(r.shape)
        s = rs / abar
        z[1:-1, 1:-1] = np.where(free, s * net[k, :nx, :ny], 0.0)
        q = r - stencil(coefs[k], z)
        rr = float((r * r).sum())
        losses.append(float((q * q).sum()) / rr)
        dz = (-2.0 / (rr * Bn)) * stencil_t(coefs[k], q)
        dnet[k, :nx, :ny] = np.where(free, s * dz[1:-1, 1:-1], 0.0)
    grad = backward(dnet, cache, bf)
    return float(np.mean(losses)), grad, losses


def _anorm_loss_and_grad(params, coefs, rs_batch, corrs, bf):
    Bn = len(rs_batch)
    feats = [features(coefs[k], rs_batch[k], bf) for k in range(Bn)]
    X0 = np.stack([f[0] for f in feats])
    net, cache = forward(params, X0, bf)
    dnet = np.zeros_like(net)
    losses = []
    for k in range(Bn):
        _, rs, abar, free = feats[k]
        r = np.asarray(rs_batch[k], np.float64)
        c = np.asarray(corrs[k], np.float64)
        nx, ny = free.shape
        s = rs / abar
        z = np.zeros(r.shape)
        z[1:-1, 1:-1] = np.where(free, s * net[k, :nx, :ny], 0.0)
        e = c - z
        ae = stencil(coefs[k], e)
        den = float((c * r).sum())
        losses.append(float((e * ae).sum()) / den)
        dz = (-1.0 / (den * Bn)) * (ae + stencil_t(coefs[k], e))
        dnet[k, :nx, :ny] = np.where(free, s * dz[1:-1, 1:-1], 0.0)
    grad = backward(dnet, cache, bf)
    return float(np.mean(losses)), grad, losses


def adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); step counts from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
