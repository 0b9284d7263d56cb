"""Closed-form checks of the oracle's pointwise functions, following the
reference's unit tests (tests/unit/*.cc): lattice tables, equilibrium and
Guo-forcing moments, BGK fixed point and linearity, Q-stencil derivatives
of polynomial fields. Each expected value is fixed by the mathematics, not
restated from the code under test."""
import numpy as np
import pytest

from lbman3d_amd import abi


@pytest.fixture(scope="module")
def po(oracle_lib):
    return oracle_lib


def test_lattice_tables(po):
    """test_lattice_stencil.cc:7-108: weights sum to one, opp is an
    involution negating the velocity, spec_axis reflects one component."""
    ex, ey, ez, w, opp, spec = po.lattice()
    e = np.stack([ex, ey, ez], axis=1)
    assert w.sum() == pytest.approx(1.0, abs=1e-15)
    assert (opp[opp] == np.arange(15)).all()
    assert (e[opp] == -e).all()
    for a in range(3):
        r = e[spec[a]]
        flipped = e.copy()
        flipped[:, a] *= -1
        assert (r == flipped).all()
    # isotropy of the D3Q15 set: sum w e_a e_b = delta_ab / 3
    np.testing.assert_allclose(np.einsum("i,ia,ib->ab", w, e, e), np.eye(3) / 3, atol=1e-15)


def params(dt=1.0, tauf=2.5):
    return abi.Params(DT=dt, TAUF=tauf).struct()


def test_feq_moments_and_zero_velocity(po):
    """test_feq_moments.cc:11-95: f_eq(rho, 0) = rho w; zeroth and first
    moments of f_eq are rho and rho u."""
    _, _, _, w, _, _ = po.lattice()
    p = params()
    feq, _, _ = po.pointwise_collide(p, 1.5, [0, 0, 0], [0, 0, 0], np.zeros(15))
    np.testing.assert_allclose(feq, 1.5 * w, rtol=1e-15)
    ex, ey, ez, *_ = po.lattice()
    rng = np.random.default_rng(0)
    for _ in range(20):
        rho, u = rng.uniform(0.8, 1.2), rng.uniform(-0.05, 0.05, 3)
        feq, _, _ = po.pointwise_collide(p, rho, u, [0, 0, 0], np.zeros(15))
        assert feq.sum() == pytest.approx(rho, abs=1e-15)
        np.testing.assert_allclose([ex @ feq, ey @ feq, ez @ feq], rho * u, atol=1e-12)


def test_guo_forcing_moments(po):
    """Guo, Zheng & Shi (2002): sum_i F_i = 0 and sum_i e_i F_i =
    (1 - DT/(2 tau)) F (the prefactor of ComputeFeqAndForcing)."""
    ex, ey, ez, *_ = po.lattice()
    rng = np.random.default_rng(1)
    for tauf in (2.5, 0.8):
        p = params(tauf=tauf)
        for _ in range(10):
            u, F = rng.uniform(-0.05, 0.05, 3), rng.uniform(-1e-3, 1e-3, 3)
            _, forcing, _ = po.pointwise_collide(p, 1.0, u, F, np.zeros(15))
            assert forcing.sum() == pytest.approx(0.0, abs=1e-17)
            np.testing.assert_allclose([ex @ forcing, ey @ forcing, ez @ forcing], (1 - 1 / (2 * tauf)) * F,
                                       rtol=1e-12, atol=1e-18)


def test_bgk_fixed_point_and_linearity(po):
    """test_bkg_collide.cc:5-57: f = f_eq with F = 0 is a fixed point; the
    forcing term is linear in F."""
    p = params()
    rng = np.random.default_rng(2)
    rho, u = 1.05, rng.uniform(-0.05, 0.05, 3)
    feq, _, _ = po.pointwise_collide(p, rho, u, [0, 0, 0], np.zeros(15))
    _, _, fstar = po.pointwise_collide(p, rho, u, [0, 0, 0], feq)
    np.testing.assert_allclose(fstar, feq, rtol=1e-14)
    F = rng.uniform(-1e-3, 1e-3, 3)
    _, f1, _ = po.pointwise_collide(p, rho, u, F, feq)
    _, f2, _ = po.pointwise_collide(p, rho, u, 2 * F, feq)
    np.testing.assert_allclose(f2, 2 * f1, rtol=1e-12, atol=1e-20)


def test_moments_half_force_correction(po):
    """test_feq_moments.cc:150-174: from f = w (rho = 1, no momentum) the
    velocity is the half-step force correction u = F DT / (2 rho)."""
    _, _, _, w, _, _ = po.lattice()
    p = params()
    F = np.array([2e-4, -1e-4, 3e-4])
    rho_u = po.moments(p, w, F)
    assert rho_u[0] == pytest.approx(1.0, abs=1e-15)
    np.testing.assert_allclose(rho_u[1:], 0.5 * F, rtol=1e-13)


def test_q_stencil_on_polynomial_fields(po):
    """test_anchored_q.cc:18-60: centred gradient of a linear field is its
    slope; the 7-point Laplacian of a quadratic is exact."""
    nx, ny, nz = 8, 7, 6
    cfg = abi.make_config(nx, ny, nz, abi.ChannelConfig)
    o = po.Oracle(cfg)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lin = (0.01 * x + 0.02 * y - 0.005 * z).reshape(-1)
    quad = (0.001 * x * x + 0.002 * y * y + 0.003 * z * z).reshape(-1)
    q = np.stack([lin, quad, lin, quad, lin])
    o.set_state(q=q)
    d = o.q_derivs(3, 3, 2)            # interior cell: rows Q components, columns (d_x, d_y, d_z, lap, ...)
    np.testing.assert_allclose(d[0, :3], [0.01, 0.02, -0.005], atol=1e-15)
    assert d[1, 3] == pytest.approx(2 * (0.001 + 0.002 + 0.003), abs=1e-14)
