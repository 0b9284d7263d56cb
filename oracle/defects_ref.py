"""Numpy restatement of the reference's plaquette winding test (test
infrastructure only: the checker of the device kernel lbm_defect_faces).

DefectFinder::ComputeWindingNumbers (src/analysis/defect_finder.tpp:12-162):
for every plaquette the director is carried around the loop of its four
cells, each next director negated when FlipN says it points against the
previous one (n1 . n2 <= 0, src/analysis/defect_fields.cc:15-24); the
plaquette is pierced when the carried director comes back against the
first. The dot products are formed as the reference writes them,
n1x*n2x + n1y*n2y + n1z*n2z, in FP64 without fused multiply-adds.
"""
from __future__ import annotations

import numpy as np


def flip_n(a, b):
    """FlipN for arrays of unit vectors [..., 3]."""
    d = a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]
    return d <= 0.0


def _loop(n1, n2, n3, n4):
    n2 = np.where(flip_n(n1, n2)[..., None], -n2, n2)
    n3 = np.where(flip_n(n2, n3)[..., None], -n3, n3)
    n4 = np.where(flip_n(n3, n4)[..., None], -n4, n4)
    return flip_n(n4, n1).astype(np.uint8)


def defect_faces(director, nx, ny, nz):
    """(def_x, def_y, def_z) flattened in the DefectFields layouts from a
    director field [V, 3] (cell = (z*ny + y)*nx + x)."""
    d = np.asarray(director, dtype=np.float64).reshape(nz, ny, nx, 3)
    # x-normal plaquettes at (x, y, z), y < ny-1, z < nz-1: (y,z) (y+1,z) (y+1,z+1) (y,z+1)
    dx = _loop(d[:-1, :-1, :], d[:-1, 1:, :], d[1:, 1:, :], d[1:, :-1, :])
    # y-normal at (x, y, z), x < nx-1, z < nz-1: (x,z) (x+1,z) (x+1,z+1) (x,z+1)
    dy = _loop(d[:-1, :, :-1], d[:-1, :, 1:], d[1:, :, 1:], d[1:, :, :-1])
    # z-normal at (x, y, z), x < nx-1, y < ny-1: (x,y) (x+1,y) (x+1,y+1) (x,y+1)
    dz = _loop(d[:, :-1, :-1], d[:, :-1, 1:], d[:, 1:, 1:], d[:, 1:, :-1])
    return dx.reshape(-1), dy.reshape(-1), dz.reshape(-1)


def reference_edges(nx, ny, nz, def_x, def_y, def_z, q_periodic=(True,) * 6):
    """The edge set of DefectFinder::BuildConnectivityGraph
    (src/analysis/defect_finder.tpp:164-340), visited as the reference visits
    it: compute_voxel (:209-304) over the interior voxels, then the six
    boundary layers with the Q-tag offsets QXoff/QYoff/QZoff (periodic wrap or
    Neumann clamp, src/offsets.h:59-83), including the `xm < nx-1` and
    `ym < ny-1` guards on the wrapped backward offsets. Returns a set of
    (smaller face id, larger face id); face ids as DefectFields::FidX/Y/Z."""
    DX, DY, DZ = (np.asarray(a, dtype=np.uint8) for a in (def_x, def_y, def_z))
    nfx, nfy = nx * (ny - 1) * (nz - 1), (nx - 1) * ny * (nz - 1)
    fx = lambda x, y, z: (z * (ny - 1) + y) * nx + x
    fy = lambda x, y, z: (z * ny + y) * (nx - 1) + x
    fz = lambda x, y, z: (z * (ny - 1) + y) * (nx - 1) + x
    IX = lambda x, y, z: fx(x, y, z)
    IY = lambda x, y, z: nfx + fy(x, y, z)
    IZ = lambda x, y, z: nfx + nfy + fz(x, y, z)
    edges = set()

    def add(a, b):
        if a != b:
            edges.add((min(a, b), max(a, b)))

    def voxel(x, y, z, xm, xp, ym, yp, zp):
        if DX[fx(x, y, z)]:                       # pt 3 and its forward neighbours, then two backward
            first = IX(x, y, z)
            for hit, other in ((DZ[fz(x, y, z)], IZ(x, y, z)), (DZ[fz(x, y, zp)], IZ(x, y, zp)),
                               (DY[fy(x, y, z)], IY(x, y, z)), (DY[fy(x, yp, z)], IY(x, yp, z)),
                               (DX[fx(xp, y, z)], IX(xp, y, z))):
                if hit:
                    add(first, other)
            if xm < nx - 1 and DZ[fz(xm, y, zp)]:
                add(first, IZ(xm, y, zp))
            if xm < nx - 1 and DY[fy(xm, yp, z)]:
                add(first, IY(xm, yp, z))
        if DY[fy(x, y, z)]:                       # pt 4, forward, then one backward
            first = IY(x, y, z)
            for hit, other in ((DZ[fz(x, y, z)], IZ(x, y, z)), (DZ[fz(x, y, zp)], IZ(x, y, zp)),
                               (DX[fx(xp, y, z)], IX(xp, y, z)), (DY[fy(x, yp, z)], IY(x, yp, z))):
                if hit:
                    add(first, other)
            if ym < ny - 1 and DZ[fz(x, ym, zp)]:
                add(first, IZ(x, ym, zp))
        if DZ[fz(x, y, z)]:                       # pt 6, forward only
            first = IZ(x, y, z)
            for hit, other in ((DX[fx(xp, y, z)], IX(xp, y, z)), (DY[fy(x, yp, z)], IY(x, yp, z)),
                               (DZ[fz(x, y, zp)], IZ(x, y, zp))):
                if hit:
                    add(first, other)

    def off(i, d, n, lo_per, hi_per):
        j = i + d
        if j < 0:
            return j + n if lo_per else 0
        if j >= n:
            return j - n if hi_per else n - 1
        return j

    p = q_periodic
    for z in range(1, nz - 2):
        for y in range(1, ny - 2):
            for x in range(1, nx - 2):
                voxel(x, y, z, x - 1, x + 1, y - 1, y + 1, z + 1)

    def bvox(x, y, z):
        voxel(x, y, z, off(x, -1, nx, p[0], p[1]), off(x, 1, nx, p[0], p[1]), off(y, -1, ny, p[2], p[3]),
              off(y, 1, ny, p[2], p[3]), off(z, 1, nz, p[4], p[5]))
    for z in sorted({0, nz - 2}):
        for y in range(ny - 1):
            for x in range(nx - 1):
                bvox(x, y, z)
    for z in range(nz - 1):
        for y in sorted({0, ny - 2}):
            for x in range(nx - 1):
                bvox(x, y, z)
    for z in range(nz - 1):
        for y in range(ny - 1):
            for x in sorted({0, nx - 2}):
                bvox(x, y, z)
    return edges
