"""HIP step against the oracle and the reference fixtures (needs a GPU).

Bar: the FP64 path is bit-identical to the reference CPU step (every fixture,
every field). The FP32 shifted-population path is compared with FP64 by
relative L2 error (tolerances stated per test). All calls go through the
C ABI of liblbman3d.so.
"""
import numpy as np
import pytest

from conftest import golden_config, golden_names, load_golden, run_golden_on
from lbman3d_amd import abi

pytestmark = pytest.mark.gpu

NAMES = golden_names()


def _solver(cfg):
    from lbman3d_amd.solver import DeviceSolver
    return DeviceSolver(cfg)


def rel_l2(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", NAMES)
def test_fp64_matches_reference_fixture_bitwise(name):
    print(f"case {name}", flush=True)
    g = load_golden(name)
    with _solver(golden_config(g)) as s:
        run_golden_on(s, g)
        st = s.state()
        np.testing.assert_array_equal(st["q"], g["q"])
        np.testing.assert_array_equal(s.force(), g["F"])
        if "rho" in g:
            np.testing.assert_array_equal(st["rho"], g["rho"])
            np.testing.assert_array_equal(st["u"], g["u"])
        if "f" in g:
            np.testing.assert_array_equal(s.populations(), g["f"])


@pytest.mark.parametrize("bc,adv", [("Channel", abi.ADV_CENTRED), ("FullyPeriodic", abi.ADV_UPWIND),
                                    ("SlitFreeSlip", abi.ADV_CENTRED), ("MovingLidZ", abi.ADV_UPWIND)])
def test_fp64_fused_step_matches_oracle(bc, adv, oracle_lib):
    """lbm_step(n) (Q kernel + fused force/collide-stream kernel) against the
    oracle's ActiveNematicSim::Step order, odd grid sizes."""
    from conftest import BC_BY_NAME
    print(f"case fused {bc} adv={adv}", flush=True)
    dims = (13, 7, 9)
    cfg = abi.make_config(*dims, BC_BY_NAME[bc], abi.Params(LAMBDA=0.5), advection=adv)
    q = abi.seed_director_field(99, *dims)
    o = oracle_lib.Oracle(cfg, threads=2)
    o.set_state(q=q)
    o.step(40)
    with _solver(cfg) as s:
        s.upload_state(q=q)
        s.step(40)
        st, ost = s.state(), o.state()
        for k in ("rho", "u", "q"):
            np.testing.assert_array_equal(st[k], ost[k])
        np.testing.assert_array_equal(s.populations(), o.populations())
        np.testing.assert_array_equal(s.force(), o.force())


def test_fp64_baseline_c0_1000_steps(oracle_lib):
    """32^3 FullyPeriodic, params.h defaults, centred, seeded random director
    (S0 = 0.5 +- 0.01), 1000 steps: GPU FP64 = reference CPU step, bitwise."""
    print("case C0 32^3 x 1000", flush=True)
    dims = (32, 32, 32)
    cfg = abi.make_config(*dims, abi.FullyPeriodicConfig, abi.Params())
    q = abi.seed_director_field(12345, *dims, 0.5, 0.01)
    o = oracle_lib.Oracle(cfg, threads=16)
    o.set_state(q=q)
    o.step(1000)
    with _solver(cfg) as s:
        s.upload_state(q=q)
        s.step(1000)
        st, ost = s.state(), o.state()
        for k in ("rho", "u", "q"):
            np.testing.assert_array_equal(st[k], ost[k])


def test_population_roundtrip_and_pull_state(oracle_lib):
    """Uploaded populations are taken as post-streaming; after steps the
    downloaded populations are the post-streaming f of the oracle."""
    print("case populations", flush=True)
    dims = (6, 5, 7)
    from conftest import BC_BY_NAME
    cfg = abi.make_config(*dims, BC_BY_NAME["MovingLidZ"], abi.Params())
    rng = np.random.default_rng(3)
    w = np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8)
    f0 = w[:, None] * (1.0 + 0.01 * rng.standard_normal((15, np.prod(dims))))
    q = abi.seed_director_field(5, *dims)
    o = oracle_lib.Oracle(cfg)
    o.set_state(q=q)
    o.set_populations(f0)
    with _solver(cfg) as s:
        s.upload_state(q=q)
        s.upload_populations(f0)
        np.testing.assert_array_equal(s.populations(), f0)
        for _ in range(3):
            s.qtensor_step()
            s.lbm_step()
            o.qtensor_step()
            o.lbm_step()
        np.testing.assert_array_equal(s.populations(), o.populations())
        np.testing.assert_array_equal(s.state()["u"], o.state()["u"])


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_SHIFTED])
def test_diagnostics_match_oracle(precision, oracle_lib):
    print(f"case diagnostics p={precision}", flush=True)
    g = load_golden("act_channel")
    cfg = golden_config(g, precision)
    o = oracle_lib.Oracle(golden_config(g))
    run_golden_on(o, g)
    with _solver(cfg) as s:
        run_golden_on(s, g)
        d = s.diagnostics()
        st = s.state()
    # Reductions of the device's own fields, FP64 accumulation.
    rho, u = st["rho"], st["u"]
    assert d["mass"] == pytest.approx(rho.sum(), rel=1e-13)
    assert d["px"] == pytest.approx((rho * u[0]).sum(), rel=1e-10, abs=1e-16)
    assert d["kinetic"] == pytest.approx(0.5 * (rho * (u * u).sum(0)).sum(), rel=1e-12)
    assert d["umax"] == pytest.approx(np.sqrt((u * u).sum(0).max()), rel=1e-12)
    assert d["nonfinite"] == 0
    ref_fe = o.free_energy()
    tol = 1e-12 if precision == abi.FP64 else 1e-5
    assert d["free_energy"] == pytest.approx(ref_fe, rel=tol)
    ls = o.log_sums()
    assert d["mass"] == pytest.approx(ls["mass"], rel=1e-12 if precision == abi.FP64 else 1e-6)


def test_order_director_matches_oracle(oracle_lib):
    print("case order/director", flush=True)
    g = load_golden("act_slit")
    o = oracle_lib.Oracle(golden_config(g))
    run_golden_on(o, g)
    oo, od = o.order_director()
    with _solver(golden_config(g)) as s:
        run_golden_on(s, g)
        so, sd = s.order_director()
    np.testing.assert_allclose(so, oo, rtol=1e-12, atol=1e-14)
    # Directors agree up to sign (n and -n are the same nematic state).
    dots = np.abs((sd * od).sum(1))
    np.testing.assert_allclose(dots, 1.0, atol=1e-9)


@pytest.mark.parametrize("name", ["act_channel", "actup_periodic", "act_slit", "act_movinglid", "unit_channel_full"])
def test_fp32_shifted_close_to_fp64(name):
    """FP32 with shifted populations against the FP64 reference fixture after
    10-20 steps: relative L2 error of Q, u and rho - rho0 below 1e-5."""
    print(f"case fp32 {name}", flush=True)
    g = load_golden(name)
    with _solver(golden_config(g, abi.FP32_SHIFTED)) as s:
        run_golden_on(s, g)
        st = s.state()
    assert rel_l2(st["q"], g["q"]) < 1e-5
    assert rel_l2(st["u"], g["u"]) < 1e-4
    rho0 = float(g["params"][1])
    assert np.max(np.abs(st["rho"] - g["rho"])) < 1e-6 * rho0


def test_errors_are_reported():
    from lbman3d_amd.solver import LbmError
    cfg = abi.make_config(4, 4, 4)
    cfg.nx = 0
    with pytest.raises(LbmError, match="grid dimensions"):
        _solver(cfg)
    # Periodic on one side of an axis only: order-dependent in the reference.
    cfg = abi.make_config(4, 4, 4)
    cfg.walls[abi.XLO].ubc = abi.U_NOSLIP
    with pytest.raises(LbmError, match="order dependent"):
        _solver(cfg)
