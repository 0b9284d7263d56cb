"""Learned closure kernels against the FP64 numpy oracle (needs a GPU).

Tolerances: the kernels hold inputs, weights, activations and
back-propagated gradients in BF16 (8-bit significand) and accumulate in
FP32. Against the oracle run with the same BF16 rounding points the forward
output agrees to relative L2 < 5e-3 and the gradient to < 2e-2 (an FP32
accumulation difference can move a stored BF16 value by one unit, 2^-8,
which then propagates); against the unrounded FP64 network to < 3e-2 and
< 6e-2.
"""
import numpy as np
import pytest

from oracle import closure_ref as cr

pytestmark = pytest.mark.gpu

GRID = (32, 8, 8)    # nx, ny, nz (nx a multiple of 16)
H = 32


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def case(seed):
    nx, ny, nz = GRID
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(8, nz * ny * nx))
    p = cr.he_init(H, seed)
    o = 0
    for ci, co in cr.layer_shapes(H):            # non-zero biases
        o += 27 * ci * co
        p[o:o + co] = rng.normal(size=co) * 0.05
        o += co
    t = rng.normal(size=(5, nz * ny * nx)) * 0.2
    return x, t, p


def grid(a):
    nx, ny, nz = GRID
    return a.reshape(a.shape[0], nz, ny, nx).transpose(1, 2, 3, 0)


def flat(g):
    return g.transpose(3, 0, 1, 2).reshape(g.shape[3], -1)


def net(seed=1):
    from lbman3d_amd.closure import ClosureNet
    return ClosureNet(*GRID, hidden=H, seed=seed)


def test_mfma_fragment_layout():
    from lbman3d_amd.solver import check, lib
    import ctypes as C
    print("case mfma layout", flush=True)
    rng = np.random.default_rng(0)
    a = cr.bf16_round(rng.normal(size=(16, 32))).astype(np.float32)
    b = cr.bf16_round(rng.normal(size=(32, 16))).astype(np.float32)
    d = np.zeros((16, 16), dtype=np.float32)
    fp = C.POINTER(C.c_float)
    check(lib().lbm_debug_mfma_bf16(a.ctypes.data_as(fp), b.ctypes.data_as(fp), d.ctypes.data_as(fp)))
    ref = a.astype(np.float64) @ b.astype(np.float64)
    np.testing.assert_allclose(d, ref, rtol=1e-5, atol=1e-5)


def test_forward_matches_oracle():
    print("case closure forward", flush=True)
    x, _, p = case(2)
    with net() as n:
        n.params = p
        np.testing.assert_array_equal(n.params, p.astype(np.float32).astype(np.float64))
        y = n.predict(x)
    ref_bf, _ = cr.forward(p, grid(x), H, bf16=True)
    ref, _ = cr.forward(p, grid(x), H)
    assert rel(y, flat(ref_bf)) < 5e-3
    assert rel(y, flat(ref)) < 3e-2


def test_gradient_matches_oracle():
    print("case closure gradient", flush=True)
    x, t, p = case(3)
    with net() as n:
        n.params = p
        L, g = n.loss_and_grad(x, t)
    Lr, gr = cr.loss_and_grad(p, grid(x), grid(t), H, bf16=True)
    L64, g64 = cr.loss_and_grad(p, grid(x), grid(t), H)
    assert L == pytest.approx(Lr, rel=2e-3)
    assert L == pytest.approx(L64, rel=2e-2)
    o = 0
    for ci, co in cr.layer_shapes(H):
        nw = 27 * ci * co
        for lo, hi in ((o, o + nw), (o + nw, o + nw + co)):
            assert rel(g[lo:hi], gr[lo:hi]) < 2e-2, (ci, co, lo)
            assert rel(g[lo:hi], g64[lo:hi]) < 6e-2, (ci, co, lo)
        o += nw + co


def test_first_adam_step():
    """After one Adam step from zero moments w1 = w0 - lr g / (|g| + eps)."""
    print("case closure adam", flush=True)
    x, t, p = case(4)
    lr = 1e-3
    with net() as n:
        n.params = p
        _, g = n.loss_and_grad(x, t)
        w0 = n.params
        n.add_sample(x, t)
        n.train(1, lr)
        w1 = n.params
    exp, _, _ = cr.adam_reference(w0, g, np.zeros_like(g), np.zeros_like(g), 1, lr)
    big = np.abs(g) > 1e-5
    np.testing.assert_allclose(w1[big], exp[big], rtol=0, atol=lr * 1e-3)
    np.testing.assert_allclose(w1, exp, rtol=0, atol=lr * 1.01)


def test_training_fits_a_teacher_network():
    print("case closure training", flush=True)
    x, _, p_teacher = case(5)
    ys = [flat(cr.forward(p_teacher, grid(xx), H)[0]) for xx in (x, x[::-1].copy())]
    xs = [x, x[::-1].copy()]
    with net(seed=9) as n:
        for xx, yy in zip(xs, ys):
            n.add_sample(xx, yy)
        first = n.train(10, 2e-3)
        for _ in range(8):
            last = n.train(40, 2e-3)
        assert np.isfinite(last)
        assert last < 0.3 * first, (first, last)
        # held-in prediction closer to the teacher than predicting zero
        assert rel(n.predict(xs[0]), ys[0]) < 0.95


def test_apply_adds_prediction_to_simulation():
    from lbman3d_amd import abi
    from lbman3d_amd.solver import DeviceSolver
    print("case closure apply", flush=True)
    nx, ny, nz = GRID
    cfg = abi.make_config(nx, ny, nz, precision=abi.FP32_SHIFTED)
    with DeviceSolver(cfg) as s, net(seed=3) as n:
        s.upload_state(q=abi.seed_director_field(5, nx, ny, nz))
        s.step(3)
        st = s.state()
        x = np.concatenate([st["q"], st["u"]])
        n.set_normalization(np.full(8, 2.0), np.full(5, 50.0))
        pred = n.predict(x)
        n.apply(s, 0.5)
        q1 = s.state()["q"]
    assert rel(q1 - st["q"], 0.5 * pred) < 1e-3


def test_rejects_bad_shapes():
    from lbman3d_amd.closure import ClosureNet
    from lbman3d_amd.solver import LbmError
    with pytest.raises(LbmError):
        ClosureNet(20, 8, 8)
    with pytest.raises(LbmError):
        ClosureNet(32, 8, 8, hidden=16)


def test_run_equals_step_and_apply_loop():
    """lbm_closure_run(n) gives the bits of n x (lbm_step(1); lbm_closure_apply)."""
    from lbman3d_amd import abi
    from lbman3d_amd.solver import DeviceSolver
    print("case closure run loop", flush=True)
    nx, ny, nz = GRID
    cfg = abi.make_config(nx, ny, nz, precision=abi.FP32_SHIFTED)
    q0 = abi.seed_director_field(6, nx, ny, nz)
    out = {}
    for use_run in (False, True):
        with DeviceSolver(cfg) as s, net(seed=3) as n:
            s.set_schedule(2)
            s.upload_state(q=q0)
            n.set_normalization(np.full(8, 2.0), np.full(5, 50.0))
            if use_run:
                ms = n.run(s, 4, 0.5)
                assert ms > 0
            else:
                for _ in range(4):
                    s.step(1)
                    n.apply(s, 0.5)
            out[use_run] = s.state()
    np.testing.assert_array_equal(out[True]["q"], out[False]["q"])
    np.testing.assert_array_equal(out[True]["u"], out[False]["u"])
