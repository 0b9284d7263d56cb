"""The fused FP64 step against the two-kernel FP64 step (needs a GPU).

On fully periodic domains the FP64 step runs one z-marching kernel that keeps
the Q-update results of the neighbouring columns on chip. It evaluates the
same per-cell arithmetic as the two unfused kernels (no FMA contraction, the
reference's operand order), so every tile shape must reproduce the two-kernel
step bit for bit: rho, u, Q, the body force and the populations, on a grid
whose sides are not multiples of any tile (partial tiles and z chunks), and
on one whose x and y sides are whole tiles (the unconditional form, every
load and store issued by every thread) with a partial last z chunk, and on
one- and two-plane grids of whole tiles, from
an uploaded state (own-slot first step) through several step calls. The
two-kernel step is itself pinned to the oracle bit for bit by the fixture
tests; the 32^3 1000-step and 256^3 tests run the fused step as the default
FP64 schedule against the oracle.
"""
import numpy as np
import pytest

from lbman3d_amd import abi

pytestmark = pytest.mark.gpu

DIMS = (37, 22, 45)
# x and y whole multiples of the 32x4, 16x8 and 32x8 tiles (their
# unconditional form) with a partial last z chunk
DIMS_WHOLE = (32, 16, 45)


def run(cfg, q, sched):
    from lbman3d_amd.solver import DeviceSolver
    with DeviceSolver(cfg) as s:
        s.set_schedule(sched)
        s.upload_state(q=q)
        s.step(5)
        s.step(7)
        st = s.state()
        st["F"] = s.force()
        st["f"] = s.populations()
        st["name"] = s.schedule_name(sched)
    return st


@pytest.mark.parametrize("dims,upwind", [(DIMS, False), (DIMS, True), (DIMS_WHOLE, True), ((32, 16, 1), True),
                                         ((32, 16, 2), False)])
def test_every_fp64_tile_equals_the_two_kernel_step(dims, upwind):
    from lbman3d_amd.solver import DeviceSolver
    cfg = abi.make_config(*dims, abi.FullyPeriodicConfig, abi.Params(ALPHA=0.046, LAMBDA=0.7 if upwind else 0.4),
                          advection=abi.ADV_UPWIND if upwind else abi.ADV_CENTRED, precision=abi.FP64)
    q = abi.seed_director_field(31, *dims)
    ref = run(cfg, q, 1)
    assert ref["name"] == "two-kernel"
    with DeviceSolver(cfg) as s:
        n = 0
        while True:
            try:
                s.schedule_name(2 + n)
            except Exception:
                break
            n += 1
    assert n >= 1
    for sched in [0] + list(range(2, 2 + n)):
        st = run(cfg, q, sched)
        print(f"case FP64 {st['name']} dims={dims} upwind={upwind}", flush=True)
        if sched >= 2:
            assert st["name"].startswith("fused:fp64:")
        for k in ("rho", "u", "q", "F", "f"):
            np.testing.assert_array_equal(st[k], ref[k], err_msg=f"{st['name']} {k}")


def test_walled_fp64_domains_keep_the_two_kernel_step():
    from lbman3d_amd.solver import DeviceSolver
    cfg = abi.make_config(12, 10, 8, abi.ChannelConfig, abi.Params(), precision=abi.FP64)
    with DeviceSolver(cfg) as s:
        assert s.schedule_name(0) == "auto:two-kernel"


def test_fp64_autotune_keeps_state_and_small_grids_pick_a_fast_schedule():
    """Tuning times every FP64 schedule on a snapshot: the state after tuning
    and stepping equals the two-kernel step's, and on a small grid (few z
    chunks to spread over the CUs) the automatic choice is not slower than
    the two-kernel split by more than timing noise."""
    from lbman3d_amd.solver import DeviceSolver
    dims = (32, 32, 32)
    cfg = abi.make_config(*dims, abi.FullyPeriodicConfig, abi.Params(), precision=abi.FP64)
    q = abi.seed_director_field(8, *dims)
    ref = run(cfg, q, 1)
    with DeviceSolver(cfg) as s:
        s.upload_state(q=q)
        res = s.autotune(3)
        print(f"case FP64 autotune 32^3: {res}", flush=True)
        assert "two-kernel" in res and len(res) >= 2
        s.step(5)
        s.step(7)
        st = s.state()
        best = min(res.values())
        assert res[s.schedule_name(0).replace("auto:", "")] == best
    for k in ("rho", "u", "q"):
        np.testing.assert_array_equal(st[k], ref[k], err_msg=k)
