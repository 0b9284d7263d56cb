"""CPU checks of the C-ABI library: it loads, exports every entry point that
include/lbman3d.h declares, and its host-only helpers (configuration,
seeded initial state, pull-slot table) behave as specified. No GPU calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT, BC_BY_NAME
from lbman3d_amd import abi

HEADER = os.path.join(ROOT, "include", "lbman3d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    from lbman3d_amd import solver
    return solver.lib()


def test_library_exports_every_declared_symbol(L):
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, f"missing exports: {missing}"
    from lbman3d_amd.solver import SIGNATURES
    assert set(names) == set(SIGNATURES), "ctypes signature table out of sync with the header"


def test_struct_layout_matches_header():
    # lbm_wall: 2 int32 + 6 doubles; lbm_config: 6 int32 + 6 walls + 11 doubles.
    assert C.sizeof(abi.LbmWall) == 8 + 6 * 8
    assert C.sizeof(abi.LbmConfig) == 24 + 6 * 56 + 11 * 8
    assert C.sizeof(abi.LbmDiag) == 9 * 8 + 8


def test_default_config_is_params_h(L):
    cfg = abi.LbmConfig()
    assert L.lbm_default_config(C.byref(cfg), 20, 100, 100) == 0
    p = cfg.p
    # src/params.h:29-112
    assert (p.dt, p.density, p.tauf, p.L, p.A, p.B, p.C) == (1.0, 1.0, 2.5, 0.06, 0.0, -0.3, 0.3)
    assert (p.lam, p.gamma, p.alpha, p.mu) == (0.3, 0.34, 0.023, 0.0)
    mine = abi.make_config(20, 100, 100)
    assert bytes(mine) == bytes(cfg)


@pytest.mark.parametrize("name", ["FullyPeriodic", "Channel", "SlitFreeSlip", "SlitNoSlip"])
def test_presets_match_python_mirror(L, name):
    cfg = abi.LbmConfig()
    L.lbm_default_config(C.byref(cfg), 8, 8, 8)
    assert L.lbm_set_preset(C.byref(cfg), name.encode()) == 0
    mine = abi.make_config(8, 8, 8, abi.PRESETS[name])
    assert bytes(mine) == bytes(cfg)
    assert L.lbm_set_preset(C.byref(cfg), b"NoSuchPreset") == abi.LBM_ERR_INVALID


def test_seeded_director_field_matches_numpy_restatement():
    from lbman3d_amd.solver import seed_director_field
    q = seed_director_field(12345, 9, 7, 5, 0.5, 0.01)
    ref = abi.seed_director_field(12345, 9, 7, 5, 0.5, 0.01)
    np.testing.assert_allclose(q, ref, rtol=0, atol=2e-15)
    # Q = S (nn - I/3): traceless, eigenvalues (2S/3, -S/3, -S/3), S in 0.5 +- 0.01
    V = q.shape[1]
    M = np.zeros((V, 3, 3))
    M[:, 0, 0], M[:, 0, 1], M[:, 0, 2] = q[0], q[1], q[2]
    M[:, 1, 0], M[:, 1, 1], M[:, 1, 2] = q[1], q[3], q[4]
    M[:, 2, 0], M[:, 2, 1], M[:, 2, 2] = q[2], q[4], -q[0] - q[3]
    ev = np.linalg.eigvalsh(M)
    S = 1.5 * ev[:, 2]
    assert np.all(np.abs(S - 0.5) <= 0.01 + 1e-12)
    np.testing.assert_allclose(ev[:, 0], -S / 3, atol=1e-12)
    np.testing.assert_allclose(ev[:, 1], -S / 3, atol=1e-12)
    # directors spread over the sphere: mean of n_z^2 ~ 1/3
    nz2 = (q[0] + q[3]) * -1.0 / S + 1.0 / 3.0
    assert abs(nz2.mean() - 1.0 / 3.0) < 0.1
    # different seeds differ, same seed repeats
    assert not np.allclose(seed_director_field(1, 9, 7, 5), q)
    np.testing.assert_array_equal(seed_director_field(12345, 9, 7, 5, 0.5, 0.01), q)


# ---- pull-slot table against a brute-force push simulation -----------------
EX = np.array([0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 1, 1, 1, 1, -1, -1, -1, -1])
OPP = [0, 3, 4, 1, 2, 6, 5, 13, 14, 11, 12, 9, 10, 7, 8]
SPEC = [[0, 3, 2, 1, 4, 5, 6, 8, 7, 10, 9, 12, 11, 14, 13],
        [0, 1, 4, 3, 2, 5, 6, 10, 9, 8, 7, 14, 13, 12, 11],
        [0, 1, 2, 3, 4, 6, 5, 11, 12, 13, 14, 7, 8, 9, 10]]


def push_writes(cfg):
    """Replay LatticeBoltzmannStep's push (src/lbm_solver.tpp:144-192) on
    tokens: returns {(dest cell, slot): (kind, src cell, src dir, face)}."""
    n = (cfg.nx, cfg.ny, cfg.nz)
    E = (EX, EY, EZ)
    out = {}
    for z in range(n[2]):
        for y in range(n[1]):
            for x in range(n[0]):
                pos = (x, y, z)
                for i in range(15):
                    d, cross = [0, 0, 0], [0, 0, 0]
                    for a in range(3):
                        t = pos[a] + E[a][i]
                        if t < 0:
                            if cfg.walls[2 * a].ubc == abi.U_PERIODIC:
                                d[a] = t % n[a]
                            else:
                                cross[a] = -1
                        elif t >= n[a]:
                            if cfg.walls[2 * a + 1].ubc == abi.U_PERIODIC:
                                d[a] = t % n[a]
                            else:
                                cross[a] = 1
                        else:
                            d[a] = t
                    if not any(cross):
                        out[(tuple(d), i)] = ("stream", pos, i, -1)
                        continue
                    for a in range(3):
                        if not cross[a]:
                            continue
                        face = 2 * a + (1 if cross[a] > 0 else 0)
                        if cfg.walls[face].ubc == abi.U_SPECULAR:
                            out[(pos, SPEC[a][i])] = ("mirror", pos, i, face)
                        else:
                            out[(pos, OPP[i])] = ("bounce", pos, i, face)
    return out


@pytest.mark.parametrize("bc", ["Channel", "FullyPeriodic", "SlitFreeSlip", "MovingLidZ", "SpecularXNoSlipY",
                                "AnchoredZ", "PoiseuilleY"])
@pytest.mark.parametrize("dims", [(4, 3, 5), (2, 3, 1), (1, 4, 3)])
def test_slot_table_reproduces_push_semantics(bc, dims):
    from lbman3d_amd.solver import slot_table
    cfg = abi.make_config(*dims, BC_BY_NAME[bc])
    tab = slot_table(cfg)
    writes = push_writes(cfg)
    n = dims
    for z in range(n[2]):
        for y in range(n[1]):
            for x in range(n[0]):
                cls = ((x == 0) | ((x == n[0] - 1) << 1)) | (((y == 0) | ((y == n[1] - 1) << 1)) << 2) | \
                      (((z == 0) | ((z == n[2] - 1) << 1)) << 4)
                for j in range(15):
                    e = int(tab[cls, j])
                    kind, src, face = e & 3, (e >> 2) & 15, (e >> 6) & 7
                    w = writes[((x, y, z), j)]
                    if kind == 0:
                        exp = ((x - EX[j]) % n[0], (y - EY[j]) % n[1], (z - EZ[j]) % n[2])
                        assert w[0] == "stream" and w[1] == exp and w[2] == j
                    else:
                        assert w[0] == ("bounce" if kind == 1 else "mirror")
                        assert w[1] == (x, y, z) and w[2] == src and w[3] == face


def test_slot_table_rejects_order_dependent_walls():
    from lbman3d_amd.solver import LbmError, slot_table
    cfg = abi.make_config(4, 4, 4)
    cfg.walls[abi.YHI].ubc = abi.U_NOSLIP      # periodic low side, wall high side
    with pytest.raises(LbmError):
        slot_table(cfg)
