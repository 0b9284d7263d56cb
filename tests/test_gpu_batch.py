"""Batched form (one persistent launch per step over many simulations)
against the same simulations stepped one by one (needs a GPU).

Each simulation of a batch runs the same fused FP32 step as a single handle,
so after a few steps the two agree to FP32 rounding (tolerance 2e-6
relative); graph replay and direct launches run identical kernels and must
agree bit for bit.
"""
import numpy as np
import pytest

from conftest import BC_BY_NAME
from lbman3d_amd import abi

pytestmark = pytest.mark.gpu

SIMS = [("FullyPeriodic", (24, 32, 20), abi.ADV_UPWIND, 0.3, 0.023),
        ("SlitFreeSlip", (32, 24, 28), abi.ADV_CENTRED, 0.5, 0.03),
        ("Channel", (20, 20, 16), abi.ADV_CENTRED, 0.4, 0.02),
        ("MovingLidZ", (16, 18, 22), abi.ADV_UPWIND, 0.6, 0.04),
        ("FullyPeriodic", (40, 16, 12), abi.ADV_CENTRED, 0.7, 0.046)]


def configs():
    return [abi.make_config(*dims, BC_BY_NAME[bc], abi.Params(LAMBDA=lam, ALPHA=zeta), advection=adv,
                            precision=abi.FP32_SHIFTED) for bc, dims, adv, lam, zeta in SIMS]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def singles(cfgs, qs, steps):
    from lbman3d_amd.solver import DeviceSolver
    out = []
    for c, q in zip(cfgs, qs):
        with DeviceSolver(c) as s:
            s.set_schedule(1)
            s.upload_state(q=q)
            s.step(steps)
            out.append(s.state())
    return out


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_single_simulations(use_graph):
    from lbman3d_amd.solver import BatchSolver, tile_count
    print(f"case batch graph={use_graph}", flush=True)
    cfgs = configs()
    qs = [abi.seed_director_field(100 + i, c.nx, c.ny, c.nz) for i, c in enumerate(cfgs)]
    ref = singles(cfgs, qs, 9)
    with BatchSolver(cfgs) as b:
        for i, q in enumerate(qs):
            b.upload_state(i, q=q)
        b.step(9, use_graph=use_graph)
        for i in range(len(cfgs)):
            st = b.state(i)
            assert rel(st["q"], ref[i]["q"]) < 2e-6, i
            assert rel(st["u"], ref[i]["u"]) < 2e-5, i
            assert np.max(np.abs(st["rho"] - ref[i]["rho"])) < 2e-6, i
            d = b.diagnostics(i)
            assert d["nonfinite"] == 0
            assert d["mass"] == pytest.approx(st["rho"].sum(), rel=1e-12)


def test_graph_replay_is_bitwise_identical_and_all_tiles_agree():
    from lbman3d_amd.solver import BatchSolver, lib
    print("case batch tiles", flush=True)
    cfgs = configs()
    qs = [abi.seed_director_field(200 + i, c.nx, c.ny, c.nz) for i, c in enumerate(cfgs)]
    states = {}
    for tile in range(lib().lbm_batch_tile_count()):
        for g in (False, True):
            with BatchSolver(cfgs) as b:
                b.set_tile(tile)
                for i, q in enumerate(qs):
                    b.upload_state(i, q=q)
                b.step(3, use_graph=g)
                b.step(8, use_graph=g)     # a second call continues the same state
                states[(tile, g)] = [b.state(i)["q"] for i in range(len(cfgs))]
    for tile in range(lib().lbm_batch_tile_count()):
        for i in range(len(cfgs)):
            np.testing.assert_array_equal(states[(tile, True)][i], states[(tile, False)][i])
            assert rel(states[(tile, True)][i], states[(0, False)][i]) < 2e-6


def test_batch_autotune_keeps_state():
    from lbman3d_amd.solver import BatchSolver
    print("case batch autotune", flush=True)
    cfgs = configs()
    qs = [abi.seed_director_field(300 + i, c.nx, c.ny, c.nz) for i, c in enumerate(cfgs)]
    ref = singles(cfgs, qs, 6)
    with BatchSolver(cfgs) as b:
        for i, q in enumerate(qs):
            b.upload_state(i, q=q)
        b.step(2)
        res = b.autotune(3)
        assert len(res) >= 2 and all(r > 0 for r in res)
        b.step(4)
        for i in range(len(cfgs)):
            assert rel(b.state(i)["q"], ref[i]["q"]) < 2e-6, i


def test_batch_rejects_fp64():
    from lbman3d_amd.solver import BatchSolver, LbmError
    c = abi.make_config(8, 8, 8, precision=abi.FP64)
    with pytest.raises(LbmError, match="FP32"):
        BatchSolver([c])


@pytest.mark.parametrize("tile_name", ["32x8x16", "p2:32x8x16"])
@pytest.mark.parametrize("limit", [2, 3])
def test_capped_grid_runs_multi_tile_and_split_paths(tile_name, limit):
    """A persistent grid of 2 or 3 blocks: every block takes many tiles (grid
    stride on the first pipeline, work queue on the third) and the third
    pipeline's final round is split into half-length items. The results
    agree with the same simulations stepped one by one, and graph replay
    equals direct launches bit for bit."""
    from lbman3d_amd.solver import BatchSolver, batch_tile_index
    print(f"case batch capped grid tile={tile_name} limit={limit}", flush=True)
    cfgs = configs()
    qs = [abi.seed_director_field(400 + i, c.nx, c.ny, c.nz) for i, c in enumerate(cfgs)]
    ref = singles(cfgs, qs, 7)
    tile = batch_tile_index(tile_name)
    out = {}
    for g in (False, True):
        with BatchSolver(cfgs) as b:
            b.set_tile(tile)
            b.set_grid_limit(limit)
            for i, q in enumerate(qs):
                b.upload_state(i, q=q)
            b.step(7, use_graph=g)
            out[g] = [b.state(i) for i in range(len(cfgs))]
    for i in range(len(cfgs)):
        for k in ("rho", "u", "q"):
            np.testing.assert_array_equal(out[True][i][k], out[False][i][k], err_msg=f"{i} {k}")
        assert rel(out[True][i]["q"], ref[i]["q"]) < 2e-6, i
        assert rel(out[True][i]["u"], ref[i]["u"]) < 2e-5, i
        assert np.max(np.abs(out[True][i]["rho"] - ref[i]["rho"])) < 2e-6, i
