"""The C++ host interface (include/lbman3d.hpp) through its check program
tests/cpp/host_api_check (built by __graft_entry__.build()).

Host mode (CPU): the reference-shaped boundary types lower to the configs
lbm_set_preset builds, and tags of another namespace map through the
traits. Run mode (GPU): DeviceFields / DeviceSolver<ChannelConfig> stepping
10 steps must give the same bits as the Python path over the same C ABI and
as the FP64 oracle.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_api_check")


def _bin():
    if not os.path.exists(BIN):
        from lbman3d_amd import build
        build.build_host_check()
    return BIN


def test_host_mode():
    r = subprocess.run([_bin(), "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout


@pytest.mark.gpu
def test_run_mode_matches_python_path_and_oracle(tmp_path):
    from lbman3d_amd import abi
    from lbman3d_amd.solver import DeviceSolver
    from oracle.pyoracle import Oracle
    print("case cpp host run", flush=True)
    out = tmp_path / "state.bin"
    r = subprocess.run([_bin(), "run", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    dims = (12, 10, 8)
    V = 12 * 10 * 8
    raw = np.fromfile(out, dtype=np.float64)
    got = {"rho": raw[:V], "u": raw[V:4 * V].reshape(3, V), "q": raw[4 * V:].reshape(5, V)}
    q = abi.seed_director_field(7, *dims, 0.5, 0.01)
    cfg = abi.make_config(*dims, abi.ChannelConfig, abi.Params())
    with DeviceSolver(cfg) as s:
        s.upload_state(q=q)
        s.step(10)
        py = s.state()
    o = Oracle(cfg)
    o.set_state(q=q)
    o.step(10)
    ref = o.state()
    for k in ("rho", "u", "q"):
        np.testing.assert_array_equal(got[k], py[k])
        np.testing.assert_array_equal(got[k], ref[k])


REFERENCE_SRC = "/root/reference/src"


@pytest.mark.skipif(not os.path.exists(os.path.join(REFERENCE_SRC, "boundary.h")),
                    reason="the reference sources are not present on this machine")
def test_binding_compiles_against_the_reference_headers(tmp_path):
    """include/lbman3d.hpp plus the five trait lines of the integration guide,
    compiled against the reference's real boundary.h, params.h and
    sim_config.h (not stand-ins): every reference preset, SimBC, the default
    BCConfig and a custom config with Anchoring and MovingWall lower to the
    configs the library builds, and params.h lowers to its defaults."""
    exe = tmp_path / "ref_boundary_check"
    src = os.path.join(ROOT, "tests", "cpp", "ref_boundary_check.cpp")
    lib_dir = os.path.join(ROOT, "lbman3d_amd")
    if not os.path.exists(os.path.join(lib_dir, "liblbman3d.so")):
        from lbman3d_amd import build
        build.build()
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-I", REFERENCE_SRC, "-I", os.path.join(ROOT, "include"), src,
           "-L", lib_dir, "-llbman3d", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
