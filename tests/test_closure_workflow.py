"""Host-side pieces of the learned-closure workflow (no GPU): the resampling
and filtering that put runs on different grids onto one measure, the
acoustic rescaling of the constants, and the equal-error interpolation."""
import numpy as np
import pytest

from lbman3d_amd import abi
from lbman3d_amd import closure as cl


def band_limited(n, kmax, seed, channels=2):
    """Real periodic fields whose modes all satisfy |k| < kmax on each axis."""
    rng = np.random.default_rng(seed)
    x = np.arange(n) * 2 * np.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    out = np.zeros((channels, n, n, n))
    for c in range(channels):
        for _ in range(6):
            k = rng.integers(-kmax + 1, kmax, size=3)
            out[c] += rng.normal() * np.cos(k[0] * X + k[1] * Y + k[2] * Z + rng.uniform(0, 2 * np.pi))
    return out.reshape(channels, n ** 3)


@pytest.mark.parametrize("n_out", [8, 12, 20, 32])
def test_fourier_resample_is_exact_on_band_limited_fields(n_out):
    a = band_limited(16, 4, n_out)
    b = cl.fourier_resample(a, 16, n_out)
    back = cl.fourier_resample(b, n_out, 16)
    np.testing.assert_allclose(back, a, atol=1e-12)
    # the mean (the k = 0 mode) is kept exactly
    np.testing.assert_allclose(b.mean(axis=1), a.mean(axis=1), atol=1e-13)


def test_fourier_resample_identity_and_sample_values():
    a = band_limited(12, 3, 1, channels=1)
    np.testing.assert_array_equal(cl.fourier_resample(a, 12, 12), a)
    # upsampling 12 -> 24 keeps the original samples on the even sites
    up = cl.fourier_resample(a, 12, 24).reshape(24, 24, 24)
    np.testing.assert_allclose(up[::2, ::2, ::2].reshape(1, -1), a, atol=1e-12)


def test_filter2_is_the_2x2x2_mean():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(3, 8 ** 3))
    f = cl.filter2(a, 8).reshape(3, 4, 4, 4)
    g = a.reshape(3, 8, 8, 8)
    assert f[1, 2, 3, 0] == pytest.approx(g[1, 4:6, 6:8, 0:2].mean(), abs=1e-15)


def test_acoustic_scaling_keeps_viscosity_and_diffusion_lengths():
    p = abi.Params()
    for r in (1.0, 8 / 7, 4 / 3, 1.6, 2.0):
        c = cl.coarse_params(p, r)
        # kinematic viscosity (tau - 1/2)/3 in lattice units scales as dt/dx^2 = 1/r
        assert (c.TAUF - 0.5) * r == pytest.approx(p.TAUF - 0.5)
        assert c.L * r * r == pytest.approx(p.L)
        assert c.GAMMA / r == pytest.approx(p.GAMMA)
        assert (c.A, c.B, c.C, c.ALPHA, c.LAMBDA) == (p.A, p.B, p.C, p.ALPHA, p.LAMBDA)
    assert cl.coarse_params(p, 1.0).__dict__ == p.__dict__


def test_zeta_draws_cover_the_stated_range_and_differ_between_sets():
    tr = cl.trajectories(200, 1000, 11)
    z = np.array([t[1] for t in tr]) / cl.ZETA_DEFAULT
    assert z.min() >= 0.5 and z.max() <= 2.0 and z.max() - z.min() > 1.3
    held = cl.trajectories(2, 2000, 12)
    assert not {t[0] for t in held} & {t[0] for t in tr}


def test_equal_error_time_interpolates_in_log_space():
    pts = [{"n": 64, "err": 4e-2, "ms": 1.0}, {"n": 96, "err": 1e-2, "ms": 4.0}, {"n": 128, "err": 0.0, "ms": 9.0}]
    assert cl.equal_error_time(pts, 2e-2) == pytest.approx(2.0)
    assert cl.equal_error_time(pts, 4e-2) == pytest.approx(1.0)
    assert cl.equal_error_time(pts, 5e-2) == pytest.approx(1.0)
    assert cl.equal_error_time(pts, 1e-3) is None
