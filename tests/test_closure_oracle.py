"""Pins of the closure network oracle (oracle/closure_ref.py), CPU only.

The forward pass is checked against a per-cell loop and against torch's
circularly padded conv3d; the gradient against central finite differences
and torch autograd in float64; BF16 rounding against torch's bfloat16
conversion; Adam against torch.optim.Adam; the box filter against an
explicit loop.
"""
import numpy as np
import pytest
import torch

from oracle import closure_ref as cr


def rand_case(shape=(4, 3, 5), hidden=4, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=shape + (cr.IN_C,))
    t = rng.normal(size=shape + (cr.OUT_C,)) * 0.1
    p = cr.he_init(hidden, seed)
    p[-cr.OUT_C:] = rng.normal(size=cr.OUT_C) * 0.1      # non-zero biases
    return x, t, p


def test_conv_matches_bruteforce_loop():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 4, 5, 3))
    w = rng.normal(size=(27, 3, 2))
    b = rng.normal(size=2)
    np.testing.assert_allclose(cr.conv(a, w, b), cr.conv_bruteforce(a, w, b), rtol=1e-13, atol=1e-13)


def test_forward_matches_torch_circular_conv3d():
    x, _, p = rand_case((5, 4, 6), hidden=6, seed=2)
    y, _ = cr.forward(p, x, 6)
    xt = torch.tensor(x, dtype=torch.float64).permute(3, 0, 1, 2)[None]     # (1, C, z, y, x)
    h = xt
    for li, (w, b) in enumerate(cr.unpack(p, 6)):
        # torch weight (cout, cin, kz, ky, kx); offset index = kz*9 + ky*3 + kx
        wt = torch.tensor(w.reshape(3, 3, 3, w.shape[1], w.shape[2])).permute(4, 3, 0, 1, 2)
        h = torch.nn.functional.conv3d(torch.nn.functional.pad(h, (1, 1, 1, 1, 1, 1), mode="circular"), wt,
                                       torch.tensor(b))
        if li < 2:
            h = torch.relu(h)
    np.testing.assert_allclose(h[0].permute(1, 2, 3, 0).numpy(), y, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    x, t, p = rand_case((4, 4, 3), hidden=4, seed=3)
    L, g = cr.loss_and_grad(p, x, t, 4)
    assert L == pytest.approx(cr.loss(p, x, t, 4), rel=1e-14)
    rng = np.random.default_rng(4)
    idx = rng.choice(len(p), 60, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (cr.loss(pp, x, t, 4) - cr.loss(pm, x, t, 4)) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-5, abs=1e-9), i


def test_gradient_matches_torch_autograd():
    x, t, p = rand_case((3, 4, 4), hidden=5, seed=5)
    _, g = cr.loss_and_grad(p, x, t, 5)
    pt = torch.tensor(p, requires_grad=True)
    xt = torch.tensor(x).permute(3, 0, 1, 2)[None]
    o = 0
    h = xt
    for li, (ci, co) in enumerate(cr.layer_shapes(5)):
        w = pt[o:o + 27 * ci * co].reshape(3, 3, 3, ci, co).permute(4, 3, 0, 1, 2)
        o += 27 * ci * co
        b = pt[o:o + co]
        o += co
        h = torch.nn.functional.conv3d(torch.nn.functional.pad(h, (1, 1, 1, 1, 1, 1), mode="circular"), w, b)
        if li < 2:
            h = torch.relu(h)
    loss = ((h[0].permute(1, 2, 3, 0) - torch.tensor(t)) ** 2).mean()
    loss.backward()
    np.testing.assert_allclose(g, pt.grad.numpy(), rtol=1e-10, atol=1e-14)


def test_a_transposed_input_gradient_is_caught():
    """A plausible mistake (input gradient without flipping the stencil)
    must fail the finite-difference pin."""
    x, t, p = rand_case((4, 3, 4), hidden=3, seed=6)
    L, g = cr.loss_and_grad(p, x, t, 3)
    # first-layer weight gradient with an unflipped input gradient of layer 2
    (w1, b1), (w2, b2), (w3, b3) = cr.unpack(p, 3)
    y, (xx, h1, h2) = cr.forward(p, x, 3)
    dy = 2.0 * (y - t) / (cr.OUT_C * x[..., 0].size)
    dh2 = sum(cr.shifted(dy, d) @ w3[k].T for k, d in enumerate(cr.OFFSETS)) * (h2 > 0)
    dh1 = sum(cr.shifted(dh2, tuple(-v for v in d)) @ w2[k].T for k, d in enumerate(cr.OFFSETS)) * (h1 > 0)
    gw1_wrong = np.stack([np.tensordot(cr.shifted(xx, d), dh1, axes=([0, 1, 2], [0, 1, 2])) for d in cr.OFFSETS])
    n1 = 27 * cr.IN_C * 3
    assert not np.allclose(gw1_wrong.ravel(), g[:n1], rtol=1e-3)


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-8, 8, 2000),
                        [1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, 65504.0, 3.0e38]])
    ref = torch.tensor(v, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(cr.bf16_round(v), ref)
    assert cr.bf16_round(np.array([1 + 2 ** -8]))[0] == 1.0                 # tie to even
    assert cr.bf16_round(np.array([1 + 3 * 2 ** -8]))[0] == 1 + 2 ** -6


def test_adam_matches_torch():
    rng = np.random.default_rng(8)
    w0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(5)]
    wt = torch.tensor(w0, requires_grad=True)
    opt = torch.optim.Adam([wt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    w, m, v = w0.copy(), np.zeros(50), np.zeros(50)
    for k, g in enumerate(grads, 1):
        opt.zero_grad()
        wt.grad = torch.tensor(g)
        opt.step()
        w, m, v = cr.adam_reference(w, g, m, v, k, 1e-2)
    np.testing.assert_allclose(w, wt.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_box_filter():
    rng = np.random.default_rng(9)
    a = rng.normal(size=(4, 6, 8, 2))
    f = cr.box_filter2(a)
    for z in range(2):
        for y in range(3):
            for x in range(4):
                np.testing.assert_allclose(f[z, y, x], a[2 * z:2 * z + 2, 2 * y:2 * y + 2, 2 * x:2 * x + 2].mean(
                    axis=(0, 1, 2)), rtol=1e-14)


def test_closure_module_filter_agrees_with_oracle_filter():
    from lbman3d_amd.closure import filter2
    rng = np.random.default_rng(10)
    n = 8
    a = rng.normal(size=(3, n ** 3))
    grid = a.reshape(3, n, n, n).transpose(1, 2, 3, 0)
    ref = cr.box_filter2(grid).transpose(3, 0, 1, 2).reshape(3, -1)
    np.testing.assert_allclose(filter2(a, n), ref, rtol=1e-14)


def test_coarse_params_acoustic_scaling():
    from lbman3d_amd import abi
    from lbman3d_amd.closure import coarse_params
    p = abi.Params()
    c = coarse_params(p)
    # same kinematic viscosity in physical units: nu dt / dx^2 halves
    assert (c.TAUF - 0.5) / 3 == pytest.approx((p.TAUF - 0.5) / 3 / 2)
    assert c.L == pytest.approx(p.L / 4) and c.GAMMA == pytest.approx(2 * p.GAMMA)
    # Q diffusivity Gamma L dt / dx^2 halves as well
    assert c.GAMMA * c.L == pytest.approx(p.GAMMA * p.L / 2)
    assert (c.A, c.B, c.C, c.ALPHA, c.LAMBDA) == (p.A, p.B, p.C, p.ALPHA, p.LAMBDA)
