"""Disclination lines of a simulation: pierced plaquettes from the device,
connectivity graph and line extraction on the host, VTK-HDF export.

Follows the reference's DefectFinder (src/analysis/defect_finder.tpp):
  1. ComputeWindingNumbers: the plaquettes a disclination pierces
     (lbm_defect_faces, computed on the GPU);
  2. BuildConnectivityGraph (:164-340): pierced plaquettes that share a
     voxel are joined, each voxel contributing the edges between its
     "forward" plaquettes and two/one backward ones, with Q-tag offsets at
     the domain faces (periodic wrap or Neumann clamp, src/offsets.h:59-83);
  3. IsolateDisclinationsFromGraph (:342-356) with the graph's FindPaths
     (src/analysis/defect_connectivity_graph.h:72-110): the edge set is
     walked into paths, each started at an odd-degree node when there is
     one, else at any node with edges, always stepping to the smallest
     neighbour (std::set order) and consuming edges;
  4. the plaquette centres become polyline points; ExportDisclinations
     (src/sim_io.cc:213-237) writes them as a VTK-HDF UnstructuredGrid of
     VTK_LINE cells (src/analysis/disclination.h:17-61).
The reference scans its candidate start nodes in std::unordered_map order,
which depends on the standard library's hashing; here the scan is in
ascending plaquette id, so the same edges are covered by the same number of
walks but the order of the lines and the starting node of a closed loop may
differ from the reference's.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi, vtkhdf
from .solver import check, lib

VTK_LINE = 3


def device_defect_faces(sim):
    """(def_x, def_y, def_z) uint8 arrays of a DeviceSolver's current state."""
    cfg = sim.cfg
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    out = [np.zeros(n, dtype=np.uint8) for n in
           (nx * (ny - 1) * (nz - 1), (nx - 1) * ny * (nz - 1), (nx - 1) * (ny - 1) * nz)]
    u8 = C.POINTER(C.c_uint8)
    check(lib().lbm_defect_faces(sim.h, *(a.ctypes.data_as(u8) for a in out)))
    return tuple(out)


def _axis_offset(i, d, n, lo_periodic, hi_periodic):
    """QAxisOffset: i + d with a periodic side wrapping and any other side
    clamping (src/offsets.h:59-83)."""
    j = i + d
    if j < 0:
        return j + n if lo_periodic else 0
    if j >= n:
        return j - n if hi_periodic else n - 1
    return j


class DefectGraph:
    """The connectivity graph of pierced plaquettes (face ids as in
    DefectFields: x faces first, then y, then z)."""

    def __init__(self, nx, ny, nz, def_x, def_y, def_z, q_periodic=(True,) * 6):
        self.n = (nx, ny, nz)
        self.nfx = nx * (ny - 1) * (nz - 1)
        self.nfy = (nx - 1) * ny * (nz - 1)
        self.dx, self.dy, self.dz = (np.asarray(a, dtype=np.uint8) for a in (def_x, def_y, def_z))
        self.qp = q_periodic
        self.adj: dict[int, set] = {}
        self._build()

    # flat indices (DefectFields::FlatX/Y/Z) and face ids (FidX/Y/Z)
    def fx(self, x, y, z):
        nx, ny, _ = self.n
        return (z * (ny - 1) + y) * nx + x

    def fy(self, x, y, z):
        nx, ny, _ = self.n
        return (z * ny + y) * (nx - 1) + x

    def fz(self, x, y, z):
        nx, ny, _ = self.n
        return (z * (ny - 1) + y) * (nx - 1) + x

    def _edge(self, a, b):
        if a == b:
            return
        self.adj.setdefault(a, set()).add(b)
        self.adj.setdefault(b, set()).add(a)

    def _voxel(self, x, y, z, xm, xp, ym, yp, zp):
        nx, ny, _ = self.n
        DX, DY, DZ = self.dx, self.dy, self.dz
        idx = lambda f: f
        iy = lambda f: self.nfx + f
        iz = lambda f: self.nfx + self.nfy + f
        if DX[self.fx(x, y, z)]:
            first = idx(self.fx(x, y, z))
            if DZ[self.fz(x, y, z)]:
                self._edge(first, iz(self.fz(x, y, z)))
            if DZ[self.fz(x, y, zp)]:
                self._edge(first, iz(self.fz(x, y, zp)))
            if DY[self.fy(x, y, z)]:
                self._edge(first, iy(self.fy(x, y, z)))
            if DY[self.fy(x, yp, z)]:
                self._edge(first, iy(self.fy(x, yp, z)))
            if DX[self.fx(xp, y, z)]:
                self._edge(first, idx(self.fx(xp, y, z)))
            if xm < nx - 1 and DZ[self.fz(xm, y, zp)]:
                self._edge(first, iz(self.fz(xm, y, zp)))
            if xm < nx - 1 and DY[self.fy(xm, yp, z)]:
                self._edge(first, iy(self.fy(xm, yp, z)))
        if DY[self.fy(x, y, z)]:
            first = iy(self.fy(x, y, z))
            if DZ[self.fz(x, y, z)]:
                self._edge(first, iz(self.fz(x, y, z)))
            if DZ[self.fz(x, y, zp)]:
                self._edge(first, iz(self.fz(x, y, zp)))
            if DX[self.fx(xp, y, z)]:
                self._edge(first, idx(self.fx(xp, y, z)))
            if DY[self.fy(x, yp, z)]:
                self._edge(first, iy(self.fy(x, yp, z)))
            if ym < ny - 1 and DZ[self.fz(x, ym, zp)]:
                self._edge(first, iz(self.fz(x, ym, zp)))
        if DZ[self.fz(x, y, z)]:
            first = iz(self.fz(x, y, z))
            if DX[self.fx(xp, y, z)]:
                self._edge(first, idx(self.fx(xp, y, z)))
            if DY[self.fy(x, yp, z)]:
                self._edge(first, iy(self.fy(x, yp, z)))
            if DZ[self.fz(x, y, zp)]:
                self._edge(first, iz(self.fz(x, y, zp)))

    def _build(self):
        nx, ny, nz = self.n
        p = self.qp
        off = lambda i, d, n, a: _axis_offset(i, d, n, p[2 * a], p[2 * a + 1])
        if not (self.dx.any() or self.dy.any() or self.dz.any()):
            return
        # interior voxels, then the six boundary layers with face offsets
        for z in range(1, nz - 2):
            for y in range(1, ny - 2):
                for x in range(1, nx - 2):
                    self._voxel(x, y, z, x - 1, x + 1, y - 1, y + 1, z + 1)

        def bvox(x, y, z):
            self._voxel(x, y, z, off(x, -1, nx, 0), off(x, 1, nx, 0), off(y, -1, ny, 1), off(y, 1, ny, 1),
                        off(z, 1, nz, 2))
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

    def paths(self):
        """FindPaths: walks consuming every edge once."""
        adj = {k: set(v) for k, v in self.adj.items()}
        out = []
        while True:
            start = None
            for node in sorted(adj):
                if len(adj[node]) % 2 == 1:
                    start = node
                    break
            if start is None:
                for node in sorted(adj):
                    if adj[node]:
                        start = node
                        break
            if start is None:
                return out
            path = [start]
            while adj[path[-1]]:
                u = path[-1]
                v = min(adj[u])
                path.append(v)
                adj[u].discard(v)
                adj[v].discard(u)
            out.append(path)

    def location(self, face):
        """Plaquette centre of a face id (DefectFields::FaceIdToLocation)."""
        nx, ny, _ = self.n
        if face >= self.nfx + self.nfy:
            f = face - self.nfx - self.nfy
            x, y, z = f % (nx - 1), (f // (nx - 1)) % (ny - 1), f // ((nx - 1) * (ny - 1))
            return (x + 0.5, y + 0.5, float(z))
        if face >= self.nfx:
            f = face - self.nfx
            x, y, z = f % (nx - 1), (f // (nx - 1)) % ny, f // ((nx - 1) * ny)
            return (x + 0.5, float(y), z + 0.5)
        x, y, z = face % nx, (face // nx) % (ny - 1), face // (nx * (ny - 1))
        return (float(x), y + 0.5, z + 0.5)

    def disclinations(self):
        """Polylines (n_i, 3) of plaquette centres, one per path."""
        return [np.array([self.location(f) for f in p], dtype=np.float64) for p in self.paths()]


def mesh(lines):
    """DisclinationMesh: Points, Connectivity, Offsets, Types of VTK_LINE
    cells joining consecutive points; lines of fewer than 2 points dropped."""
    pts, conn, offs, types = [], [], [0], []
    npts = 0
    for ln in lines:
        k = len(ln)
        if k < 2:
            continue
        pts.append(np.asarray(ln, dtype=np.float64).reshape(k, 3))
        for i in range(k - 1):
            conn += [npts + i, npts + i + 1]
            offs.append(offs[-1] + 2)
        types += [VTK_LINE] * (k - 1)
        npts += k
    points = np.concatenate(pts) if pts else np.zeros((0, 3))
    return points, np.array(conn, dtype=np.int64), np.array(offs, dtype=np.int64), np.array(types, dtype=np.uint8)


def unstructured_tree(points, connectivity, offsets, types) -> vtkhdf.H5File:
    """UnstructuredGridWriter::WriteTopology (src/vtkhdf_writer.h:380-410)."""
    f = vtkhdf.H5File()
    root = f.root.group("VTKHDF", {"Type": "UnstructuredGrid", "Version": np.array([2, 0], dtype=np.int64)})
    root.dataset("NumberOfPoints", np.array([len(points)], dtype=np.int64))
    root.dataset("NumberOfCells", np.array([len(types)], dtype=np.int64))
    root.dataset("NumberOfConnectivityIds", np.array([len(connectivity)], dtype=np.int64))
    root.dataset("Points", np.asarray(points, dtype=np.float64).reshape(-1, 3))
    root.dataset("Connectivity", np.asarray(connectivity, dtype=np.int64))
    root.dataset("Offsets", np.asarray(offsets, dtype=np.int64))
    root.dataset("Types", np.asarray(types, dtype=np.uint8))
    return f


def find_disclinations(sim):
    """Disclination polylines of a DeviceSolver's current state."""
    cfg = sim.cfg
    faces = device_defect_faces(sim)
    qp = tuple(cfg.walls[f].qbc == abi.Q_PERIODIC for f in range(6))
    return DefectGraph(cfg.nx, cfg.ny, cfg.nz, *faces, q_periodic=qp).disclinations()


def export_disclinations(sim, directory: str, frame: int, num_steps: int) -> str:
    """Write ``<directory>/disclinations_<frame>.vtkhdf``."""
    tree = unstructured_tree(*mesh(find_disclinations(sim)))
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, f"disclinations_{frame:0{vtkhdf.step_width(num_steps)}d}.vtkhdf")
    tree.save(path)
    return path
