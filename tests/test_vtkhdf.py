"""VTK-HDF frames (lbman3d_amd/vtkhdf.py) in the reference's layout
(SimIO::ExportVTKHDF, src/sim_io.cc:168-211; ImageDataWriter,
src/vtkhdf_writer.h:115-160).

The files are checked two ways: with the module's own reader, and with the
HDF Group's h5dump (libhdf5 1.10) where the image has one (/opt/conda/bin):
every attribute's type, shape and value, every dataset's type and shape,
and every dataset's bytes.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from lbman3d_amd import vtkhdf

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def frame(nx=6, ny=5, nz=4, seed=0):
    V = nx * ny * nz
    rng = np.random.default_rng(seed)
    fields = {k: rng.normal(size=V) for k in ("rho", "ux", "uy", "uz", "order")}
    fields["director"] = rng.normal(size=(V, 3))
    return fields


def write(tmp_path, nx=6, ny=5, nz=4, seed=0, debug=False):
    f = frame(nx, ny, nz, seed)
    pops = None
    if debug:
        pops = np.random.default_rng(seed + 1).normal(size=(15, nx * ny * nz))
        f.update({f"f{i}": pops[i] for i in range(15)})
    tree = vtkhdf.image_data_tree(nx, ny, nz, f["rho"], f["ux"], f["uy"], f["uz"], f["order"], f["director"],
                                  pops)
    path = tmp_path / "lbm_0.vtkhdf"
    tree.save(path)
    return path, f


def test_round_trip_with_own_reader(tmp_path):
    nx, ny, nz = 6, 5, 4
    path, f = write(tmp_path, nx, ny, nz)
    t = vtkhdf.read_h5(path)
    v = t["children"]["VTKHDF"]
    a = v["attrs"]
    assert a["Type"] == "ImageData"
    np.testing.assert_array_equal(a["Version"], [2, 0])
    np.testing.assert_array_equal(a["WholeExtent"], [0, nx - 1, 0, ny - 1, 0, nz - 1])
    np.testing.assert_array_equal(a["Origin"], [0, 0, 0])
    np.testing.assert_array_equal(a["Spacing"], [1, 1, 1])
    np.testing.assert_array_equal(a["Direction"], np.eye(3).reshape(9))
    pd = v["children"]["PointData"]["children"]
    assert set(pd) == {"rho", "order", "ux", "uy", "uz", "director"}
    for k in ("rho", "order", "ux", "uy", "uz"):
        assert pd[k]["data"].shape == (nz, ny, nx)
        np.testing.assert_array_equal(pd[k]["data"].reshape(-1), f[k])
    assert pd["director"]["data"].shape == (nz, ny, nx, 3)
    np.testing.assert_array_equal(pd["director"]["data"].reshape(-1, 3), f["director"])


def test_step_width_matches_reference_rule():
    assert vtkhdf.step_width(1) == 1 and vtkhdf.step_width(10) == 1 and vtkhdf.step_width(11) == 2
    assert vtkhdf.step_width(100001) == 6


def test_debug_frame_holds_the_populations(tmp_path):
    nx, ny, nz = 5, 4, 3
    path, f = write(tmp_path, nx, ny, nz, debug=True)
    pd = vtkhdf.read_h5(path)["children"]["VTKHDF"]["children"]["PointData"]["children"]
    assert set(pd) == {"rho", "order", "ux", "uy", "uz", "director"} | {f"f{i}" for i in range(15)}
    for i in range(15):
        assert pd[f"f{i}"]["data"].shape == (nz, ny, nx)
        np.testing.assert_array_equal(pd[f"f{i}"]["data"].reshape(-1), f[f"f{i}"])


@pytest.mark.skipif(H5DUMP is None, reason="no h5dump in this image")
@pytest.mark.parametrize("debug", [False, True])
def test_libhdf5_reads_the_frame(tmp_path, debug):
    nx, ny, nz = 7, 5, 3
    path, f = write(tmp_path, nx, ny, nz, seed=3, debug=debug)
    names = ["rho", "order", "ux", "uy", "uz"] + ([f"f{i}" for i in range(15)] if debug else [])
    r = subprocess.run([H5DUMP, "-H", str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    hdr = r.stdout
    for name in ("Direction", "Origin", "Spacing", "Type", "Version", "WholeExtent"):
        assert f'ATTRIBUTE "{name}"' in hdr
    assert "STRSIZE 9;" in hdr and "H5T_STR_NULLPAD" in hdr and "H5T_CSET_ASCII" in hdr
    for name in names:
        assert re.search(rf'DATASET "{name}" {{\s*DATATYPE  H5T_IEEE_F64LE\s*DATASPACE  SIMPLE '
                         rf'{{ \( {nz}, {ny}, {nx} \) / \( {nz}, {ny}, {nx} \) }}', hdr), name
    assert re.search(rf'DATASET "director" {{\s*DATATYPE  H5T_IEEE_F64LE\s*DATASPACE  SIMPLE '
                     rf'{{ \( {nz}, {ny}, {nx}, 3 \)', hdr)
    r = subprocess.run([H5DUMP, "-a", "/VTKHDF/Version", "-a", "/VTKHDF/WholeExtent", "-a", "/VTKHDF/Type",
                        "-a", "/VTKHDF/Direction", str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "(0): 2, 0" in r.stdout
    assert f"(0): 0, {nx - 1}, 0, {ny - 1}, 0, {nz - 1}" in r.stdout
    assert '(0): "ImageData"' in r.stdout
    assert "(0): 1, 0, 0, 0, 1, 0, 0, 0, 1" in r.stdout
    for name in names + ["director"]:
        out = tmp_path / f"{name}.bin"
        r = subprocess.run([H5DUMP, "-d", f"/VTKHDF/PointData/{name}", "-b", "LE", "-o", str(out), str(path)],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stderr
        np.testing.assert_array_equal(np.fromfile(out, dtype="<f8"), np.asarray(f[name]).reshape(-1))


@pytest.mark.gpu
def test_export_frame_from_device(tmp_path):
    from lbman3d_amd import abi
    from lbman3d_amd.solver import DeviceSolver
    print("case vtkhdf export", flush=True)
    dims = (12, 10, 8)
    cfg = abi.make_config(*dims, abi.ChannelConfig)
    with DeviceSolver(cfg) as s:
        s.upload_state(q=abi.seed_director_field(9, *dims))
        s.step(5)
        path = vtkhdf.export_frame(s, str(tmp_path / "data"), 3, 1000)
        dpath = vtkhdf.export_frame(s, str(tmp_path / "debug"), 3, 1000, debug=True)
        st = s.state()
        order, director = s.order_director()
        pops = s.populations()
    assert os.path.basename(path) == "lbm_003.vtkhdf"
    pd = vtkhdf.read_h5(path)["children"]["VTKHDF"]["children"]["PointData"]["children"]
    np.testing.assert_array_equal(pd["rho"]["data"].reshape(-1), st["rho"])
    np.testing.assert_array_equal(pd["ux"]["data"].reshape(-1), st["u"][0])
    np.testing.assert_array_equal(pd["order"]["data"].reshape(-1), order)
    np.testing.assert_array_equal(pd["director"]["data"].reshape(-1, 3), director)
    dpd = vtkhdf.read_h5(dpath)["children"]["VTKHDF"]["children"]["PointData"]["children"]
    for i in range(15):
        np.testing.assert_array_equal(dpd[f"f{i}"]["data"].reshape(-1), pops[i])
    np.testing.assert_array_equal(dpd["rho"]["data"].reshape(-1), st["rho"])
