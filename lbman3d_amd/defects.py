"""Disclination lines of a simulation: pierced plaquettes from the device,
connectivity graph and line extraction on the host, VTK-HDF export.

Follows the reference's DefectFinder (src/analysis/defect_finder.tpp):
  1. ComputeWindingNumbers: the plaquettes a disclination pierces
     (lbm_defect_faces, computed on the GPU);
  2. BuildConnectivityGraph (:164-340): pierced plaquettes that share a
     voxel are joined (the edge set is stated per voxel in DefectGraph);
  3. IsolateDisclinationsFromGraph (:342-356) with the graph's FindPaths
     (src/analysis/defect_connectivity_graph.h:72-110): the edge set is
     walked into paths, each started at an odd-degree node when there is
     one, else at any node with edges, always stepping to the smallest
     neighbour (std::set order) and consuming edges;
  4. the plaquette centres become polyline points; ExportDisclinations
     (src/sim_io.cc:213-237) writes them as a VTK-HDF UnstructuredGrid of
     VTK_LINE cells (src/analysis/disclination.h:17-61).
The reference scans its candidate start nodes in std::unordered_map order,
which depends on the standard library's hash table; here the scan is in
ascending plaquette id. Both cover every edge exactly once. A line or loop
without junctions (every plaquette of degree <= 2) becomes the same walk up
to its direction and, for a loop, its starting plaquette. Where lines meet
(degree > 2) the split into walks depends on the start order, and the number
of walks can differ from the reference's.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import vtkhdf
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


class DefectGraph:
    """The connectivity graph of pierced plaquettes (face ids as in
    DefectFields: x faces first, then y, then z).

    Edges, by voxel: a voxel (x, y, z), 0 <= x < nx-1 (likewise y, z), has
    six plaquette faces, a low and a high one along each axis. Every two
    pierced faces of one voxel are joined, except that the pairs of high
    faces (x-high with y-high, x-high with z-high) are joined only where
    x < nx-2, and the pair (y-high, z-high) only where y < ny-2. This is the
    edge set that BuildConnectivityGraph (src/analysis/defect_finder.tpp:
    164-340) produces: it adds the high-high pairs of a voxel when it visits
    the next voxel along x or y, which does not exist in the last layer, and
    it visits no voxel across a periodic seam (its wrapped backward offsets
    are skipped by the `xm < nx-1` and `ym < ny-1` guards). The graph
    therefore does not depend on the Q boundary tags."""

    # the 15 face pairs of a voxel; faces 0..5 = x-lo, x-hi, y-lo, y-hi, z-lo, z-hi
    _PAIRS = [(a, b) for a in range(6) for b in range(a + 1, 6)]

    def __init__(self, nx, ny, nz, def_x, def_y, def_z):
        self.n = (nx, ny, nz)
        self.nfx = nx * (ny - 1) * (nz - 1)
        self.nfy = (nx - 1) * ny * (nz - 1)
        self.dx, self.dy, self.dz = (np.asarray(a, dtype=np.uint8) for a in (def_x, def_y, def_z))
        self.adj: dict[int, set] = {}
        self._build()

    # flat indices (DefectFields::FlatX/Y/Z)
    def fx(self, x, y, z):
        nx, ny, _ = self.n
        return (z * (ny - 1) + y) * nx + x

    def fy(self, x, y, z):
        nx, ny, _ = self.n
        return (z * ny + y) * (nx - 1) + x

    def fz(self, x, y, z):
        nx, ny, _ = self.n
        return (z * (ny - 1) + y) * (nx - 1) + x

    def edges(self) -> set:
        """Every edge once, as (smaller face id, larger face id)."""
        return {(a, b) for a, nb in self.adj.items() for b in nb if a < b}

    def _build(self):
        nx, ny, nz = self.n
        if min(nx, ny, nz) < 2:
            return
        DX = self.dx.reshape(nz - 1, ny - 1, nx).astype(bool)
        DY = self.dy.reshape(nz - 1, ny, nx - 1).astype(bool)
        DZ = self.dz.reshape(nz, ny - 1, nx - 1).astype(bool)
        ix = np.arange(self.nfx).reshape(nz - 1, ny - 1, nx)
        iy = self.nfx + np.arange(self.nfy).reshape(nz - 1, ny, nx - 1)
        iz = self.nfx + self.nfy + np.arange((nx - 1) * (ny - 1) * nz).reshape(nz, ny - 1, nx - 1)
        # the six faces of every voxel, as (pierced, id) arrays of shape (nz-1, ny-1, nx-1)
        faces = [(DX[:, :, :-1], ix[:, :, :-1]), (DX[:, :, 1:], ix[:, :, 1:]),
                 (DY[:, :-1, :], iy[:, :-1, :]), (DY[:, 1:, :], iy[:, 1:, :]),
                 (DZ[:-1], iz[:-1]), (DZ[1:], iz[1:])]
        last_x = np.zeros((nz - 1, ny - 1, nx - 1), dtype=bool)
        last_x[:, :, -1] = True
        last_y = np.zeros_like(last_x)
        last_y[:, -1, :] = True
        skip = {(1, 3): last_x, (1, 5): last_x, (3, 5): last_y}
        for a, b in self._PAIRS:
            m = faces[a][0] & faces[b][0]
            if (a, b) in skip:
                m = m & ~skip[(a, b)]
            for u, v in zip(faces[a][1][m].tolist(), faces[b][1][m].tolist()):
                self.adj.setdefault(u, set()).add(v)
                self.adj.setdefault(v, set()).add(u)

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
    return DefectGraph(cfg.nx, cfg.ny, cfg.nz, *faces).disclinations()


def export_disclinations(sim, directory: str, frame: int, num_steps: int) -> str:
    """Write ``<directory>/disclinations_<frame>.vtkhdf``."""
    tree = unstructured_tree(*mesh(find_disclinations(sim)))
    os.makedirs(directory, exist_ok=True)
    path = os.path.join(directory, f"disclinations_{frame:0{vtkhdf.step_width(num_steps)}d}.vtkhdf")
    tree.save(path)
    return path
