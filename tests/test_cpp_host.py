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
