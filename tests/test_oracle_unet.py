"""Pins the U-Net oracle (oracle/unet.py) before it judges the HIP kernels.

* forward and parameter gradients against torch CPU autograd in FP64
  (conv2d / avg_pool2d / nearest interpolate / cat: library routines);
* the loss gradient against central finite differences;
* bf16 rounding against torch's bfloat16 cast; Adam against torch.optim.Adam;
* A^T against the scipy CSR transpose;
* closed forms: zero parameters give z = 0 and loss 1; M(c r) = c M(r) for
  c > 0 and M for the matrix c A is M / c (the feature normalisation).
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import krylov, unet
from unet_systems import small_system


def rand_params(seed, scale=1.0):
    p = unet.init_params(seed).astype(np.float64)
    rng = np.random.default_rng(seed + 100)
    sl, _ = unet.param_slices()
    for k, (o, s) in sl.items():
        if k.endswith(".b"):
            p[o:o + int(np.prod(s))] = 0.1 * rng.standard_normal(int(np.prod(s)))
    return p * scale


def torch_forward(params, X0):
    P = {k: torch.tensor(v) for k, v in unet.unpack(params).items()}
    x = torch.tensor(X0).permute(0, 3, 1, 2)

    def conv(x, n):
        return torch.relu(F.conv2d(x, P[n + ".w"].permute(0, 3, 1, 2), P[n + ".b"], padding=1))

    def up(x):
        return F.interpolate(x, scale_factor=2, mode="nearest")

    e0 = conv(x, "e0")
    e1 = conv(F.avg_pool2d(e0, 2), "e1")
    e2 = conv(F.avg_pool2d(e1, 2), "e2")
    e3 = conv(F.avg_pool2d(e2, 2), "e3")
    d2 = conv(torch.cat([up(e3), e2], 1), "d2")
    d1 = conv(torch.cat([up(d2), e1], 1), "d1")
    d0 = conv(torch.cat([up(d1), e0], 1), "d0")
    net = torch.einsum("bchw,c->bhw", d0, P["out.w"]) + P["out.b"][0]
    return net, P


def test_param_count():
    assert unet.param_count() == 448 + 4640 + 18496 + 73856 + 110656 + 27680 + 6928 + 17


def test_forward_and_backward_match_torch_autograd():
    rng = np.random.default_rng(0)
    X0 = rng.standard_normal((2, 16, 24, 3))
    params = rand_params(1)
    net, cache = unet.forward(params, X0, bf=False)
    tnet, _ = torch_forward(params, X0)
    np.testing.assert_allclose(net, tnet.detach().numpy(), rtol=1e-12, atol=1e-12)
    dnet = rng.standard_normal(net.shape)
    # autograd through a fresh graph with leaf parameters
    leaves = {k: torch.tensor(v, requires_grad=True) for k, v in unet.unpack(params).items()}
    x = torch.tensor(X0).permute(0, 3, 1, 2)

    def conv(x, n):
        return torch.relu(F.conv2d(x, leaves[n + ".w"].permute(0, 3, 1, 2), leaves[n + ".b"], padding=1))

    up = lambda t: F.interpolate(t, scale_factor=2, mode="nearest")
    e0 = conv(x, "e0"); e1 = conv(F.avg_pool2d(e0, 2), "e1"); e2 = conv(F.avg_pool2d(e1, 2), "e2")
    e3 = conv(F.avg_pool2d(e2, 2), "e3")
    d2 = conv(torch.cat([up(e3), e2], 1), "d2"); d1 = conv(torch.cat([up(d2), e1], 1), "d1")
    d0 = conv(torch.cat([up(d1), e0], 1), "d0")
    tn = torch.einsum("bchw,c->bhw", d0, leaves["out.w"]) + leaves["out.b"][0]
    (tn * torch.tensor(dnet)).sum().backward()
    g = unet.unpack(unet.backward(dnet, cache, bf=False))
    for k, t in leaves.items():
        np.testing.assert_allclose(g[k], t.grad.numpy(), rtol=1e-10, atol=1e-10, err_msg=k)


def test_stencil_transpose_matches_csr():
    coef, _ = small_system(1)
    coef = coef + np.random.default_rng(2).standard_normal(coef.shape) * (coef != 0)   # not symmetric
    y = np.random.default_rng(3).standard_normal(coef.shape[1:])
    A = krylov.to_csr(coef)
    np.testing.assert_allclose(unet.stencil_t(coef, y).ravel(), A.T @ y.ravel(), rtol=1e-13, atol=1e-13)


def test_zero_params_give_zero_output_and_unit_loss():
    coef, r = small_system(4)
    z = unet.apply(np.zeros(unet.param_count()), coef, r)
    assert not z.any()
    loss, grad, _ = unet.loss_and_grad(np.zeros(unet.param_count()), coef[None], r[None], bf=False)
    assert loss == pytest.approx(1.0, rel=1e-15)
    # with all activations zero only the output bias has a gradient
    g = unet.unpack(grad)
    assert g["out.b"][0] != 0 and all(not v.any() for k, v in g.items() if k != "out.b")


def test_output_scaling_laws():
    coef, r = small_system(5)
    p = rand_params(6, 0.5)
    z = unet.apply(p, coef, r, bf=False)
    assert np.abs(z).max() > 0
    np.testing.assert_allclose(unet.apply(p, coef, 3.0 * r, bf=False), 3.0 * z, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(unet.apply(p, 4.0 * coef, r, bf=False), z / 4.0, rtol=1e-12, atol=1e-14)
    # identity rows and the halo stay zero
    free = (coef[1] != 0) | (coef[2] != 0) | (coef[3] != 0) | (coef[4] != 0)
    assert not z[~free].any()


def test_loss_gradient_finite_differences():
    coefs, rs = zip(*(small_system(s) for s in (7, 8)))
    coefs, rs = np.stack(coefs), np.stack(rs)
    p = rand_params(9, 0.7)
    loss, grad, _ = unet.loss_and_grad(p, coefs, rs, bf=False)
    sl, n = unet.param_slices()
    rng = np.random.default_rng(10)
    idx = [sl[k][0] + int(rng.integers(int(np.prod(sl[k][1])))) for k in sl]
    h = 1e-6
    for k in idx:
        e = np.zeros(n); e[k] = h
        lp = unet.loss_and_grad(p + e, coefs, rs, bf=False)[0]
        lm = unet.loss_and_grad(p - e, coefs, rs, bf=False)[0]
        fd = (lp - lm) / (2 * h)
        assert fd == pytest.approx(grad[k], rel=1e-5, abs=1e-9), k


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(11)
    a = np.concatenate([rng.standard_normal(10000) * 10.0 ** rng.integers(-30, 30, 10000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, 1e-40]]).astype(np.float32)
    want = torch.tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(unet.bf16(a), want)


def test_adam_matches_torch():
    rng = np.random.default_rng(12)
    p0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(5)]
    tp = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, 1):
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = unet.adam(p, g, m, v, k)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-13, atol=1e-15)


def test_anorm_loss_gradient_and_value():
    """A-norm error loss: finite differences for the gradient; the value is
    ||c - z||_A^2 / ||c||_A^2 with c the exact solution (scipy CSR)."""
    from scipy.sparse.linalg import spsolve
    systems = [small_system(s) for s in (13, 14)]
    coefs = np.stack([c for c, _ in systems])
    rs = np.stack([r for _, r in systems])
    corrs = np.stack([krylov.direct_solve(c, r) for c, r in systems])
    p = rand_params(15, 0.7)
    loss, grad, losses = unet.loss_and_grad(p, coefs, rs, bf=False, corrs=corrs)
    for k in range(2):
        A = krylov.to_csr(coefs[k])
        z = unet.apply(p, coefs[k], rs[k], bf=False).ravel()
        c = corrs[k].ravel()
        e = c - z
        assert losses[k] == pytest.approx(float(e @ (A @ e)) / float(c @ (A @ c)), rel=1e-10)
    sl, n = unet.param_slices()
    rng = np.random.default_rng(16)
    h = 1e-6
    for key in ("e0.w", "e3.b", "d2.w", "d0.b", "out.w", "out.b"):
        o, s = sl[key]
        k = o + int(rng.integers(int(np.prod(s))))
        e = np.zeros(n); e[k] = h
        fd = (unet.loss_and_grad(p + e, coefs, rs, bf=False, corrs=corrs)[0]
              - unet.loss_and_grad(p - e, coefs, rs, bf=False, corrs=corrs)[0]) / (2 * h)
        assert fd == pytest.approx(grad[k], rel=1e-5, abs=1e-9), key
