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
