"""Write the golden fixtures of tests/golden/*.npz from the reference's own
CPU solver core (oracle/_ref/ref_<scenario>, built from /root/reference/src by
oracle/build_ref.sh). Runs only where the reference sources exist; the
fixtures (inputs + outputs, raw FP64) are what travels.

Each fixture records the scenario's physical constants as the reference's
params.h files define them (cited per scenario), the boundary preset, the
stepping mode, the input state and the outputs after `steps` steps:
rho, u, Q, F, f (post-streaming, i*V + cell), the free energy and the
SimIO::Log sums.

The disclination fixture (defects_ref.npz) holds two Q fields on a
14 x 12 x 10 grid and, per boundary preset, the reference finder's order,
director, pierced plaquettes and disclination paths.

usage: python tests/golden/make_golden.py [defects]
"""
from __future__ import annotations

import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from lbman3d_amd.abi import seed_director_field, uniform_stream  # noqa: E402
from oracle.pyoracle import run_ref  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# Physical constants of each compiled scenario (dt, density, tauf, L, A, B, C,
# lambda, gamma, alpha, mu) and grid.
_DEFAULT = dict(dt=1.0, density=1.0, tauf=2.5 * 1.0, L=0.06, A=0.0, B=-0.3, C=0.3, lam=0.3, gamma=0.34,
                alpha=0.023, mu=0.0)
SCENARIOS = {
    # src/params.h defaults on 12x10x8
    "act": (dict(_DEFAULT), (12, 10, 8), "centred"),
    # src/params.h with LAMBDA 0.7, ALPHA 0.046 (2x default), upwind: the 256^3 workload's constants
    "actup": (dict(_DEFAULT, lam=0.7, alpha=0.046), (12, 10, 8), "upwind"),
    # tests/params/unit/params.h: DT 0.05, TAUF 1.5*DT, L 0.01, ALPHA 0.04
    "unit": (dict(_DEFAULT, dt=0.05, tauf=1.5 * 0.05, L=0.01, alpha=0.04), (8, 8, 8), "centred"),
    # tests/params/coupled/params.h: TAUF 1.5, L 0.01, ALPHA 0
    "coupled": (dict(_DEFAULT, tauf=1.5, L=0.01, alpha=0.0), (16, 16, 16), "centred"),
    # tests/params/qtrelax/params.h: TAUF 1.5, L 0.01, ALPHA 0
    "qtrelax": (dict(_DEFAULT, tauf=1.5, L=0.01, alpha=0.0), (16, 16, 16), "centred"),
    # tests/params/bc_check/params.h: 8x8x16, TAUF 1.5, L 0.3, ALPHA 0
    "bccheck": (dict(_DEFAULT, tauf=1.5, L=0.3, alpha=0.0), (8, 8, 16), "centred"),
    # tests/params/poiseuille/params.h: 5x16x5, TAUF sqrt(3/16)+0.5, L 0.01
    "poiseuille": (dict(_DEFAULT, tauf=math.sqrt(3.0 / 16.0) + 0.5, L=0.01, alpha=0.04), (5, 16, 5), "centred"),
}
# Poiseuille driving force kDeltaP = 8 nu u_max / ny^2, nu = (2 TAUF - 1)/6
_NU = (2 * SCENARIOS["poiseuille"][0]["tauf"] - 1) / 6.0
KDELTAP = 8 * _NU * 0.1 / float(16 * 16)

# (name, scenario, bc, mode, force_model, steps, init)
CASES = [
    ("act_channel", "act", "Channel", "full", 0, 20, "director"),
    ("act_periodic", "act", "FullyPeriodic", "full", 0, 20, "director"),
    ("act_slit", "act", "SlitFreeSlip", "full", 0, 20, "director"),
    ("act_slitnoslip", "act", "SlitNoSlip", "full", 0, 20, "director"),
    ("act_anchored", "act", "AnchoredZ", "full", 0, 12, "director"),
    ("act_movinglid", "act", "MovingLidZ", "full", 0, 12, "director"),
    ("act_specx_noslipy", "act", "SpecularXNoSlipY", "full", 0, 12, "director"),
    ("actup_periodic", "actup", "FullyPeriodic", "full", 0, 20, "director"),
    ("actup_channel", "actup", "Channel", "full", 0, 20, "director"),
    ("unit_specz_lbm", "unit", "SpecularZ", "lbm", 0, 30, "force"),
    ("unit_channel_full", "unit", "Channel", "full", 0, 10, "director"),
    ("qtrelax_q", "qtrelax", "FullyPeriodic", "q", 1, 20, "distorted"),
    ("coupled_full", "coupled", "FullyPeriodic", "full", 0, 15, "distorted"),
    ("bccheck_anchored_q", "bccheck", "AnchoredZ", "q", 1, 20, "distorted"),
    ("poiseuille_lbm", "poiseuille", "PoiseuilleY", "lbm", 0, 200, "poiseuille"),
]


def distorted_q(seed, dims, noise=0.05):
    """The reference tests' distorted initial Q shape (Qxx 0.33, Qyy -0.15,
    uniform noise of amplitude NOISE on every component,
    tests/integration/test_qtensor_relaxation.cc:21-31), drawn from the
    counter-based stream instead of std::mt19937."""
    V = int(np.prod(dims))
    u = [uniform_stream(seed, V, 8 + k) for k in range(5)]
    base = [0.33, 0.0, 0.0, -0.15, 0.0]
    return np.stack([base[k] + noise * (2 * u[k] - 1) for k in range(5)])


def make_case(name, scen, bc, mode, fm, steps, init, seed):
    prm, dims, adv = SCENARIOS[scen]
    V = int(np.prod(dims))
    fields = force = None
    if init == "director":
        q = seed_director_field(seed, *dims, 0.5, 0.01)
    elif init == "distorted":
        q = distorted_q(seed, dims)
    else:
        q = np.zeros((5, V))
    if init == "force":
        force = np.stack([np.full(V, 1e-5), np.zeros(V), np.zeros(V)])
    if init == "poiseuille":
        force = np.stack([np.full(V, KDELTAP), np.zeros(V), np.zeros(V)])
    r = run_ref(scen, bc, mode, fm, steps, dims, q, fields=fields, force=force)
    out = dict(
        scenario=scen, bc=bc, mode=mode, force_model=fm, steps=steps, dims=np.array(dims),
        advection=adv, params=np.array([prm[k] for k in ("dt", "density", "tauf", "L", "A", "B", "C", "lam",
                                                         "gamma", "alpha", "mu")]),
        q0=q, has_force0=force is not None, q=r["q"], F=r["F"], free_energy=r["free_energy"], log=r["log"])
    if force is not None:
        out["force0"] = force
    if mode != "q":   # a Q-only run never touches the fluid fields
        out.update(rho=r["rho"], u=r["u"])
        if V <= 1024:  # populations only for the small grids (fixture size)
            out["f"] = r["f"]
    np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **out)
    print(f"{name}: {r['stderr'].strip()}")


# ---- disclination finder (oracle/_ref/ref_defects: QtensorToOrderDirector and
# DefectFinder<BC> of the reference, src/analysis_fields.cc and
# src/analysis/defect_finder.tpp, on a 14 x 12 x 10 grid) ----
DEFECT_DIMS = (14, 12, 10)


def q_of_director(d, S=0.5):
    """Q = S (n n - I/3) of unit directors d (V x 3), as 5 x V."""
    return np.stack([S * (d[:, 0] ** 2 - 1 / 3), S * d[:, 0] * d[:, 1], S * d[:, 0] * d[:, 2],
                     S * (d[:, 1] ** 2 - 1 / 3), S * d[:, 1] * d[:, 2]])


def wedge_lines_q(dims, seed):
    """Two straight wedge lines (+1/2 along z through plaquette centre
    (3.5, 4.5), -1/2 along x through (y, z) = (7.5, 5.5)), superposed as
    angles, with a constant out-of-plane tilt (no director lies in the xy
    plane, where the reference's eigenvector formula degenerates) and a
    little seeded noise."""
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    phi = 0.5 * np.arctan2(y - 4.5, x - 3.5)
    psi = -0.5 * np.arctan2(z - 5.5, y - 7.5)
    rng = np.random.default_rng(seed)
    th = (phi + psi).reshape(-1) + 0.05 * rng.standard_normal(nx * ny * nz)
    c = np.sqrt(1 - 0.35 ** 2)
    d = np.stack([c * np.cos(th), c * np.sin(th), np.full_like(th, 0.35)], axis=1)
    return q_of_director(d)


def make_defects():
    from oracle.pyoracle import run_ref_defects
    dims = DEFECT_DIMS
    out = dict(dims=np.array(dims))
    fields = {"random": seed_director_field(77, *dims, 0.5, 0.01), "lines": wedge_lines_q(dims, 5)}
    for name, q in fields.items():
        out[f"{name}_q"] = q
        for bc in ("FullyPeriodic", "Channel"):
            r = run_ref_defects(bc, dims, q)
            key = f"{name}_{bc}"
            out[f"{key}_order"] = r["order"]
            out[f"{key}_director"] = r["director"]
            for a, f in zip("xyz", r["faces"]):
                out[f"{key}_def_{a}"] = f
            out[f"{key}_path_len"] = np.array([len(p) for p in r["paths"]], dtype=np.int64)
            out[f"{key}_path_points"] = (np.concatenate(r["paths"]) if r["paths"] else np.zeros((0, 3)))
            print(f"defects {key}: {sum(int(f.sum()) for f in r['faces'])} pierced plaquettes, "
                  f"{len(r['paths'])} paths")
    np.savez_compressed(os.path.join(OUT, "defects_ref.npz"), **out)


def main():
    for k, case in enumerate(CASES):
        make_case(*case, seed=1000 + k)
    make_defects()


if __name__ == "__main__":
    if sys.argv[1:] == ["defects"]:
        make_defects()
    else:
        main()
