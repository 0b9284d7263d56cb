"""Shared test setup: the `gpu` marker, repository import path, and helpers
that turn a golden fixture back into a configuration."""
from __future__ import annotations

import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from lbman3d_amd import abi  # noqa: E402

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X); runs the HIP library")


# Boundary presets by the names the reference harness uses.
_Pe, _Ne, _NS, _Sp = abi.Periodic(), abi.Neumann(), abi.NoSlip(), abi.SpecularReflection()
_PP = abi.WallSpec(_Pe, _Pe)
BC_BY_NAME = {
    "FullyPeriodic": abi.FullyPeriodicConfig,
    "Channel": abi.ChannelConfig,
    "SlitFreeSlip": abi.SlitConfig,
    "SlitNoSlip": abi.SlitNoSlipConfig,
    "PoiseuilleY": abi.BCConfig(_PP, _PP, abi.WallSpec(_Ne, _NS), abi.WallSpec(_Ne, _NS), _PP, _PP, "PoiseuilleY"),
    "SpecularZ": abi.BCConfig(_PP, _PP, _PP, _PP, abi.WallSpec(_Ne, _Sp), abi.WallSpec(_Ne, _Sp), "SpecularZ"),
    "AnchoredZ": abi.BCConfig(_PP, _PP, _PP, _PP,
                              abi.WallSpec(abi.Anchoring(0.5, np.pi / 4.0, 0.0), _NS),
                              abi.WallSpec(abi.Anchoring(0.5, np.pi / 4.0, 0.0), _NS), "AnchoredZ"),
    "MovingLidZ": abi.BCConfig(_PP, _PP, _PP, _PP, abi.WallSpec(_Ne, _NS),
                               abi.WallSpec(abi.Anchoring(0.4, 0.3, 1.1), abi.MovingWall(0.02, -0.01, 0.0)),
                               "MovingLidZ"),
    "SpecularXNoSlipY": abi.BCConfig(abi.WallSpec(_Ne, _Sp), abi.WallSpec(_Ne, _Sp), abi.WallSpec(_Ne, _NS),
                                     abi.WallSpec(_Ne, _NS), _PP, _PP, "SpecularXNoSlipY"),
}


def golden_names():
    """The solver fixtures (the disclination fixture, defects_ref.npz, has
    its own test)."""
    names = (os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(GOLDEN_DIR, "*.npz")))
    return sorted(n for n in names if n != "defects_ref")


def load_golden(name):
    with np.load(os.path.join(GOLDEN_DIR, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def golden_config(g, precision=abi.FP64):
    p = g["params"]
    params = abi.Params(DT=p[0], kDensity=p[1], TAUF=p[2], L=p[3], A=p[4], B=p[5], C=p[6], LAMBDA=p[7],
                        GAMMA=p[8], ALPHA=p[9], MU=p[10])
    adv = abi.ADV_UPWIND if str(g["advection"]) == "upwind" else abi.ADV_CENTRED
    dims = [int(d) for d in g["dims"]]
    return abi.make_config(*dims, BC_BY_NAME[str(g["bc"])], params, advection=adv, precision=precision,
                           force_model=int(g["force_model"]))


def run_golden_on(sim, g):
    """Drive an object with the solver interface (oracle or device) through
    the fixture's protocol: initial state, optional force, then steps."""
    sim.set_state(q=g["q0"]) if hasattr(sim, "set_state") else sim.upload_state(q=g["q0"])
    if bool(g["has_force0"]):
        F = g["force0"]
        sim.set_force(F) if hasattr(sim, "set_force") else sim.upload_force(F)
    mode = str(g["mode"])
    for _ in range(int(g["steps"])):
        if mode in ("full", "q"):
            sim.qtensor_step()
        if mode in ("full", "lbm"):
            sim.lbm_step()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle
