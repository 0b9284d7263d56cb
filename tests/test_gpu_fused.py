"""The fused single-kernel FP32 step against the two-kernel FP32 split and
against the FP64 reference (needs a GPU).

Both FP32 schedules evaluate the same pointwise formulas; they differ only in
FMA contraction, in where intermediates live and, in the third pipeline, in
the summation order of the body force (divergence form), so after a few
steps they agree to FP32 rounding (tolerance 2e-6 relative, stated per
test). Odd grid
sizes exercise partial tiles and partial z chunks; every boundary type the
reference has is covered.
"""
import numpy as np
import pytest

from conftest import BC_BY_NAME, golden_config, load_golden
from lbman3d_amd import abi

pytestmark = pytest.mark.gpu

CASES = [("FullyPeriodic", (37, 29, 23), abi.ADV_UPWIND), ("Channel", (40, 20, 12), abi.ADV_CENTRED),
         ("SlitFreeSlip", (24, 30, 40), abi.ADV_CENTRED), ("MovingLidZ", (20, 18, 35), abi.ADV_UPWIND),
         ("AnchoredZ", (16, 16, 16), abi.ADV_CENTRED), ("SpecularXNoSlipY", (19, 21, 9), abi.ADV_UPWIND),
         # cells far from the walls (periodic specialisation) next to the wall shells
         ("SlitNoSlip", (80, 36, 24), abi.ADV_CENTRED), ("Channel", (72, 44, 20), abi.ADV_UPWIND),
         ("MovingLidZ", (24, 20, 48), abi.ADV_UPWIND),
         # overlapping shells: every cell within two cells of a wall
         ("Channel", (3, 5, 10), abi.ADV_CENTRED),
         # thin in z: one and two planes (the z march wraps onto the same planes)
         ("FullyPeriodic", (40, 24, 2), abi.ADV_UPWIND), ("Channel", (24, 20, 1), abi.ADV_CENTRED),
         # whole tiles in x and y: the schedules whose memory instructions are
         # unconditional run only on such grids
         ("FullyPeriodic", (128, 64, 40), abi.ADV_UPWIND)]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def run(cfg, q, schedule, steps):
    from lbman3d_amd.solver import DeviceSolver
    with DeviceSolver(cfg) as s:
        s.set_schedule(schedule)
        s.upload_state(q=q)
        s.step(steps)
        st = s.state()
        st["f"] = s.populations()
        st["F"] = s.force()
    return st


def tiles():
    from lbman3d_amd.solver import tile_count
    return list(range(2, 2 + tile_count()))


@pytest.mark.parametrize("bc,dims,adv", CASES)
def test_fused_matches_two_kernel_all_tiles(bc, dims, adv):
    cfg = abi.make_config(*dims, BC_BY_NAME[bc], abi.Params(LAMBDA=0.5, ALPHA=0.03), advection=adv,
                          precision=abi.FP32_SHIFTED)
    q = abi.seed_director_field(31, *dims, 0.5, 0.01)
    ref = run(cfg, q, 1, 8)
    assert np.isfinite(ref["q"]).all()
    for sch in tiles():
        print(f"case {bc} {dims} schedule {sch}", flush=True)
        st = run(cfg, q, sch, 8)
        assert rel(st["q"], ref["q"]) < 2e-6, sch
        assert rel(st["u"], ref["u"]) < 2e-5, sch
        assert np.max(np.abs(st["rho"] - ref["rho"])) < 2e-6, sch
        assert rel(st["f"], ref["f"]) < 2e-6, sch
        assert rel(st["F"], ref["F"]) < 2e-4, sch


@pytest.mark.parametrize("name", ["act_channel", "actup_periodic", "act_slit", "act_movinglid", "act_anchored"])
def test_fused_fp32_close_to_fp64_reference(name):
    """lbm_step(n) on the auto (fused) schedule against the reference FP64
    fixture: relative L2 of Q below 1e-5, of u below 1e-4."""
    print(f"case fused-vs-fp64 {name}", flush=True)
    from lbman3d_amd.solver import DeviceSolver
    g = load_golden(name)
    assert str(g["mode"]) == "full"
    with DeviceSolver(golden_config(g, abi.FP32_SHIFTED)) as s:
        s.upload_state(q=g["q0"])
        s.step(int(g["steps"]))
        st = s.state()
    l2 = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert l2(st["q"], g["q"]) < 1e-5
    assert l2(st["u"], g["u"]) < 1e-4


def test_autotune_leaves_state_unchanged():
    print("case autotune", flush=True)
    from lbman3d_amd.solver import DeviceSolver
    dims = (48, 40, 36)
    cfg = abi.make_config(*dims, abi.FullyPeriodicConfig, abi.Params(), precision=abi.FP32_SHIFTED)
    q = abi.seed_director_field(8, *dims)
    with DeviceSolver(cfg) as s:
        s.upload_state(q=q)
        s.step(3)
        before = s.state()
        fb = s.populations()
        res = s.autotune(3)
        assert len(res) >= 2 and all(v > 0 for v in res.values())
        after = s.state()
        np.testing.assert_array_equal(after["q"], before["q"])
        np.testing.assert_array_equal(after["u"], before["u"])
        np.testing.assert_array_equal(s.populations(), fb)
        s.step(2)
    with DeviceSolver(cfg) as s2:
        s2.upload_state(q=q)
        s2.set_schedule(1)
        s2.step(5)
        ref = s2.state()
    with DeviceSolver(cfg) as s3:
        s3.upload_state(q=q)
        s3.step(3)
        s3.autotune(3)
        s3.step(2)
        assert rel(s3.state()["q"], ref["q"]) < 2e-6


def test_fused_large_periodic_grid():
    """128^3, the benchmark's boundary case at a size where many blocks run."""
    print("case 128^3", flush=True)
    n = 128
    cfg = abi.make_config(n, n, n, abi.FullyPeriodicConfig, abi.Params(ALPHA=0.046, LAMBDA=0.7),
                          advection=abi.ADV_UPWIND, precision=abi.FP32_SHIFTED)
    q = abi.seed_director_field(2, n, n, n)
    a = run(cfg, q, 1, 5)
    b = run(cfg, q, 2, 5)
    assert rel(b["q"], a["q"]) < 2e-6
    assert rel(b["f"], a["f"]) < 2e-6
