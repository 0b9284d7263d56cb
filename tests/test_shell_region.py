"""The wall-shell cell enumeration of the library (host side, no GPU).

A step on a grid with walls runs the periodic specialisation over every tile
and then two one-thread-per-cell passes over the cells within 3 (Q update)
and 2 (force and collide-stream) cells of an open face. Those passes must
visit every such cell exactly once and nothing else; here the library's own
enumeration is compared with a brute-force scan of the grid for every
boundary preset on thin and ragged grids.
"""
import ctypes
import itertools

import numpy as np
import pytest

from conftest import BC_BY_NAME
from lbman3d_amd import abi


def open_faces(cfg):
    return [not (cfg.walls[f].qbc == abi.Q_PERIODIC and cfg.walls[f].ubc == abi.U_PERIODIC) for f in range(6)]


def brute_force(dims, faces, w):
    cells = set()
    for p in itertools.product(*(range(n) for n in dims)):
        if any((faces[2 * a] and p[a] < w) or (faces[2 * a + 1] and p[a] >= dims[a] - w) for a in range(3)):
            cells.add(p)
    return cells


@pytest.mark.parametrize("bc", sorted(BC_BY_NAME))
def test_shell_region_matches_brute_force(bc):
    from lbman3d_amd.solver import LbmError, shell_region
    checked = 0
    for dims in [(1, 5, 4), (3, 5, 10), (4, 4, 4), (5, 3, 2), (7, 9, 6), (19, 21, 9), (24, 20, 1), (33, 6, 5)]:
        cfg = abi.make_config(*dims, BC_BY_NAME[bc], precision=abi.FP32_SHIFTED)
        faces = open_faces(cfg)
        for w in (2, 3):
            try:
                got = shell_region(cfg, w)
            except LbmError:
                continue   # wall combination the library rejects on this grid
            cells = [tuple(int(v) for v in p) for p in got]
            assert all(0 <= c[a] < dims[a] for c in cells for a in range(3)), (bc, dims, w)
            assert len(cells) == len(set(cells)), (bc, dims, w)
            assert set(cells) == brute_force(dims, faces, w), (bc, dims, w)
            if w == 2:
                # the x band comes first, then y outside it, then z outside both
                order = [0 if (faces[0] and c[0] < 2) or (faces[1] and c[0] >= dims[0] - 2)
                         else 1 if (faces[2] and c[1] < 2) or (faces[3] and c[1] >= dims[1] - 2) else 2
                         for c in cells]
                assert order == sorted(order), (bc, dims)
            checked += 1
    if not any(open_faces(abi.make_config(4, 4, 4, BC_BY_NAME[bc]))):
        return
    assert checked > 0


def test_periodic_grid_has_no_shell():
    from lbman3d_amd.solver import shell_region
    cfg = abi.make_config(6, 5, 4, abi.FullyPeriodicConfig, precision=abi.FP32_SHIFTED)
    assert shell_region(cfg, 2).shape == (0, 3)
    assert shell_region(cfg, 3).shape == (0, 3)


def test_tile_map_stays_inside_the_domain():
    """Every plane the fused pipelines request (z0 - 2 .. z1 + 1 of a chunk)
    and every tile column (any distance) maps into [0, n): periodic sides
    wrap, other sides clamp. One-plane grids were the edge case."""
    from lbman3d_amd.solver import lib
    L = lib()
    out = ctypes.c_int32()
    for n in range(1, 7):
        for lo in (0, 1):
            for hi in (0, 1):
                for kind, ts in ((0, range(-2, n + 2)), (1, range(-3 * n - 5, 4 * n + 6))):
                    for t in ts:
                        assert L.lbm_debug_tile_map(kind, t, n, lo, hi, ctypes.byref(out)) == 0
                        if t < 0:
                            want = t % n if lo else 0
                        elif t >= n:
                            want = t % n if hi else n - 1
                        else:
                            want = t
                        assert out.value == want, (kind, t, n, lo, hi)
