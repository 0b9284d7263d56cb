"""Disclination finding (lbm_defect_faces on the device, lbman3d_amd/defects.py
on the host) against the reference's DefectFinder
(src/analysis/defect_finder.tpp), pinned by director fields whose
disclinations are known in closed form.

The winding test's numpy restatement (oracle/defects_ref.py) is checked on a
uniform field, straight +-1/2 wedge lines along each axis, an integer
(+1) defect and a loop; the device kernel must then give the same bytes as
the restatement on the device's own director field.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from lbman3d_amd import abi, defects
from oracle import defects_ref as dr

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def wedge(nx, ny, nz, k, axis="z", centre=None):
    """Director of a straight wedge line of charge k along `axis` through
    `centre` (a plaquette centre): n = (cos k phi, sin k phi) in the plane
    normal to the axis."""
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    a, b = {"z": (x, y), "x": (y, z), "y": (z, x)}[axis]
    ca, cb = centre
    phi = np.arctan2(b - cb, a - ca)
    d = np.zeros((nz, ny, nx, 3))
    ia, ib = {"z": (0, 1), "x": (1, 2), "y": (2, 0)}[axis]
    d[..., ia] = np.cos(k * phi)
    d[..., ib] = np.sin(k * phi)
    return d.reshape(-1, 3)


def test_uniform_field_has_no_defects():
    d = np.tile([0.3, -0.2, 0.9], (6 * 5 * 4, 1))
    assert all(not f.any() for f in dr.defect_faces(d, 6, 5, 4))


@pytest.mark.parametrize("k", [0.5, -0.5])
def test_straight_half_line_along_z(k):
    nx, ny, nz = 8, 7, 6
    fx, fy, fz = dr.defect_faces(wedge(nx, ny, nz, k, "z", (3.5, 2.5)), nx, ny, nz)
    assert not fx.any() and not fy.any()
    fz = fz.reshape(nz, ny - 1, nx - 1)
    expect = np.zeros_like(fz)
    expect[:, 2, 3] = 1
    np.testing.assert_array_equal(fz, expect)


def test_straight_half_line_along_x_and_y():
    nx, ny, nz = 6, 7, 8
    fx, fy, fz = dr.defect_faces(wedge(nx, ny, nz, 0.5, "x", (2.5, 3.5)), nx, ny, nz)   # (y, z) centre
    assert not fy.any() and not fz.any()
    fx = fx.reshape(nz - 1, ny - 1, nx)
    assert fx[3, 2, :].all() and fx.sum() == nx
    fx, fy, fz = dr.defect_faces(wedge(nx, ny, nz, -0.5, "y", (4.5, 1.5)), nx, ny, nz)  # (z, x) centre
    assert not fx.any() and not fz.any()
    fy = fy.reshape(nz - 1, ny, nx - 1)
    assert fy[4, :, 1].all() and fy.sum() == ny


def test_escaped_integer_defect_is_not_a_disclination():
    """A +1 radial defect escaped into z (the director tilts to z at the
    core) is topologically trivial: no plaquette is pierced, unlike the
    half-integer lines above."""
    nx, ny, nz = 10, 10, 3
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    rx, ry = x - 4.5, y - 4.5
    r, phi = np.hypot(rx, ry), np.arctan2(ry, rx)
    theta = 0.5 * np.pi * np.minimum(r / 4.0, 1.0)
    d = np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], axis=-1)
    assert not any(f.any() for f in dr.defect_faces(d.reshape(-1, 3), nx, ny, nz))


def test_graph_traces_a_straight_line():
    nx, ny, nz = 8, 7, 6
    faces = dr.defect_faces(wedge(nx, ny, nz, 0.5, "z", (3.5, 2.5)), nx, ny, nz)
    g = defects.DefectGraph(nx, ny, nz, *faces)
    lines = g.disclinations()
    assert len(lines) == 1
    np.testing.assert_array_equal(lines[0], [[3.5, 2.5, float(z)] for z in range(nz)])


def test_graph_closes_a_loop():
    """A small ring through the voxels (2,2,2) (3,2,2) (3,3,2) (2,3,2): each
    voxel has two pierced faces, so the graph is one 4-cycle and the walk
    returns to its start."""
    nx, ny, nz = 6, 6, 6
    dx = np.zeros(nx * (ny - 1) * (nz - 1), np.uint8)
    dy = np.zeros((nx - 1) * ny * (nz - 1), np.uint8)
    dz = np.zeros((nx - 1) * (ny - 1) * nz, np.uint8)
    g0 = defects.DefectGraph(nx, ny, nz, dx, dy, dz)
    dx[g0.fx(3, 2, 2)] = dx[g0.fx(3, 3, 2)] = 1
    dy[g0.fy(3, 3, 2)] = dy[g0.fy(2, 3, 2)] = 1
    g = defects.DefectGraph(nx, ny, nz, dx, dy, dz)
    paths = g.paths()
    assert len(paths) == 1 and paths[0][0] == paths[0][-1] and len(paths[0]) == 5
    assert all(len(v) == 2 for v in g.adj.values()) and len(g.adj) == 4


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 4), (5, 4, 3), (6, 7, 5), (9, 8, 7)])
@pytest.mark.parametrize("density", [0.08, 0.3, 0.7])
def test_voxel_pair_graph_equals_the_reference_visit(dims, density):
    """The graph built per voxel (pairs of pierced faces) has exactly the
    edges of the reference's visit order (oracle restatement of
    BuildConnectivityGraph) for every combination of periodic and clamped Q
    faces, on random pierced sets from sparse to dense."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 100 + ny * 10 + nz + int(density * 1000))
    sizes = (nx * (ny - 1) * (nz - 1), (nx - 1) * ny * (nz - 1), (nx - 1) * (ny - 1) * nz)
    faces = [(rng.random(n) < density).astype(np.uint8) for n in sizes]
    got = defects.DefectGraph(nx, ny, nz, *faces).edges()
    for mask in range(8):
        qp = tuple(bool(mask >> a & 1) for a in range(3) for _ in range(2))
        assert got == dr.reference_edges(nx, ny, nz, *faces, q_periodic=qp), qp


def test_voxel_pair_graph_last_layer_rule_is_material():
    """Joining the high-high face pairs of the last voxel layer too (plain
    all-pairs per voxel) would add an edge the reference never adds."""
    nx, ny, nz = 4, 4, 4
    sizes = (nx * (ny - 1) * (nz - 1), (nx - 1) * ny * (nz - 1), (nx - 1) * (ny - 1) * nz)
    faces = [np.ones(n, dtype=np.uint8) for n in sizes]
    g = defects.DefectGraph(nx, ny, nz, *faces)
    ref = dr.reference_edges(nx, ny, nz, *faces)
    x_hi, y_hi = g.fx(nx - 1, 0, 0), g.nfx + g.fy(nx - 2, 1, 0)     # faces of voxel (nx-2, 0, 0)
    assert (min(x_hi, y_hi), max(x_hi, y_hi)) not in ref
    assert g.edges() == ref


def test_mesh_and_unstructured_export(tmp_path):
    lines = [np.array([[0.5, 0.5, 0.0], [0.5, 0.5, 1.0], [0.5, 0.5, 2.0]]), np.array([[1.0, 2.5, 0.5]]),
             np.array([[2.5, 1.0, 3.5], [3.5, 1.0, 3.5]])]
    pts, conn, offs, types = defects.mesh(lines)
    assert pts.shape == (5, 3)
    np.testing.assert_array_equal(conn, [0, 1, 1, 2, 3, 4])
    np.testing.assert_array_equal(offs, [0, 2, 4, 6])
    np.testing.assert_array_equal(types, [3, 3, 3])
    path = tmp_path / "disclinations_0.vtkhdf"
    defects.unstructured_tree(pts, conn, offs, types).save(path)
    t = defects.vtkhdf.read_h5(path)["children"]["VTKHDF"]
    assert t["attrs"]["Type"] == "UnstructuredGrid"
    np.testing.assert_array_equal(t["children"]["Points"]["data"], pts)
    np.testing.assert_array_equal(t["children"]["NumberOfCells"]["data"], [3])
    if H5DUMP:
        r = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        assert 'DATASET "Types"' in r.stdout and "H5T_STD_U8LE" in r.stdout
        assert 'DATASET "Points"' in r.stdout and "( 5, 3 )" in r.stdout


@pytest.mark.gpu
def test_device_faces_match_restatement():
    from lbman3d_amd.solver import DeviceSolver
    print("case defect faces", flush=True)
    for dims, bc in (((24, 20, 16), abi.FullyPeriodicConfig), ((20, 16, 12), abi.ChannelConfig)):
        cfg = abi.make_config(*dims, bc, precision=abi.FP32_SHIFTED)
        with DeviceSolver(cfg) as s:
            s.upload_state(q=abi.seed_director_field(3, *dims))
            s.step(20)
            faces = defects.device_defect_faces(s)
            _, director = s.order_director()
            lines = defects.find_disclinations(s)
        ref = dr.defect_faces(director, *dims)
        for a, b in zip(faces, ref):
            np.testing.assert_array_equal(a, b)
        assert sum(int(f.sum()) for f in faces) > 0 and len(lines) > 0


@pytest.mark.gpu
def test_device_finds_a_seeded_wedge_line(tmp_path):
    """A +1/2 wedge line along x uploaded as Q = S (n n - I/3). The line is
    placed so that no cell has n_z = 0: the reference's eigenvector formula
    (QtensorToOrderDirector) degenerates for directors in the xy plane, and
    the director is taken from Q on the device exactly as it does."""
    from lbman3d_amd.solver import DeviceSolver
    print("case defect wedge", flush=True)
    nx, ny, nz = 8, 16, 8
    d = wedge(nx, ny, nz, 0.5, "x", (7.5, 3.5))          # centre (y, z)
    S = 0.5
    q = np.stack([S * (d[:, 0] ** 2 - 1 / 3), S * d[:, 0] * d[:, 1], S * d[:, 0] * d[:, 2],
                  S * (d[:, 1] ** 2 - 1 / 3), S * d[:, 1] * d[:, 2]])
    with DeviceSolver(abi.make_config(nx, ny, nz)) as s:
        s.upload_state(q=q)
        lines = defects.find_disclinations(s)
        path = defects.export_disclinations(s, str(tmp_path), 0, 1)
    assert len(lines) == 1
    np.testing.assert_array_equal(lines[0], [[float(x), 7.5, 3.5] for x in range(nx)])
    assert os.path.basename(path) == "disclinations_0.vtkhdf"
