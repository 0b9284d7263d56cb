"""lbm.log writer (lbman3d_amd/simio.py) against the reference's SimIO
(src/sim_io.cc:19-165).

The number format is pinned against std::to_chars of this image's g++ (the
definition std::format("{}") uses for a double) on random and edge values;
the line layout against the reference's format strings; the divergence
verdict against its NaN rule.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from lbman3d_amd import abi
from lbman3d_amd.simio import SimLog, fmt_double

TOCHARS = r'''
#include <charconv>
#include <cstdio>
#include <cstring>
int main() {
    double v;
    char buf[64];
    while (std::fread(&v, sizeof v, 1, stdin) == 1) {
        auto r = std::to_chars(buf, buf + sizeof buf, v);
        *r.ptr = 0;
        std::puts(buf);
    }
}
'''


@pytest.fixture(scope="module")
def to_chars(tmp_path_factory):
    d = tmp_path_factory.mktemp("tochars")
    src, exe = d / "t.cpp", d / "t"
    src.write_text(TOCHARS)
    r = subprocess.run(["g++", "-std=c++17", "-O1", str(src), "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.skip(f"no floating-point std::to_chars: {r.stderr[-200:]}")

    def run(vals):
        arr = np.asarray(vals, dtype=np.float64)
        out = subprocess.run([str(exe)], input=arr.tobytes(), capture_output=True, check=True)
        return out.stdout.decode().split("\n")[:len(arr)]
    return run


def test_fmt_double_matches_to_chars(to_chars):
    rng = np.random.default_rng(3)
    vals = list(rng.normal(size=300) * 10.0 ** rng.integers(-30, 30, 300))
    vals += list(rng.integers(-10 ** 6, 10 ** 6, 50).astype(float))
    vals += [0.0, -0.0, 1.0, 960.0, 1e-5, 1e-4, 1e-3, 0.1, 1 / 3, 1e15, 1e16, 1e17, 123456789012.0, 5e-324,
             1.7976931348623157e308, 2.5, 0.023, 0.34, 2.0 / 3.0, 4096.0, 1e21, 1.5e-7, 100.0, 12345.678,
             math.inf, -math.inf]
    got = [fmt_double(v) for v in vals]
    ref = to_chars(vals)
    bad = [(v, g, r) for v, g, r in zip(vals, got, ref) if g != r]
    assert not bad, bad[:10]


def test_log_lines_and_divergence(tmp_path):
    cfg = abi.make_config(20, 100, 100, abi.ChannelConfig)
    path = tmp_path / "lbm.log"
    d = {"mass": 200000.0, "px": 1.5e-3, "py": -2e-4, "pz": 0.0, "kinetic": 1e-6, "free_energy": -12.5,
         "du2": 0.0, "u2": 0.0, "umax": 0.0}
    with SimLog(str(path)) as log:
        log.setup_summary(cfg, "Channel", "HIP device 0: test")
        assert log.log(d, 0)
        d2 = dict(d, du2=1e-8, u2=4e-6, umax=0.01)
        assert log.log(d2, 1000)
        assert not log.log(dict(d, mass=math.nan), 2000)
    text = path.read_text().splitlines()
    assert text[0] == "Hybrid Lattice Boltzmann simulation for 3D active nematics"
    assert "  nx = 20, ny = 100, nz = 100" in text
    assert "  BC = Channel" in text
    assert "  TAUF                = 2.5" in text
    assert "  Kinematic Viscosity = 0.6666666666666666" in text     # kDensity/3 * (TAUF - DT/2)
    assert "  omega               = 0.6" in text
    assert "  ALPHA (numerical)  = 0.023" in text
    line0 = [t for t in text if t.startswith("Time 0:")][0]
    assert line0 == ("Time 0: Mass: 2e+05, Px: 0.0015, Py: -2e-04, Pz: 0, Kinetic Energy: 1e-06, "
                     "Total Energy: -12.499999, Relative Error: -nan, NumDisclinations: 0, "
                     "Nematic Energy: -12.5, Max |u|: 0")
    line1 = [t for t in text if t.startswith("Time 1000:")][0]
    assert "Relative Error: 0.0025," in line1 and "Max |u|: 0.01" in line1
    assert text[-2] == "DIVERGED at time step 2000 — aborting."
    assert text[-1] == "LBM program exited."


@pytest.mark.gpu
def test_run_loop_logs_device_diagnostics(tmp_path, oracle_lib):
    """The main-loop driver on the device: every logged line carries the
    oracle's SimIO sums for the same step (FP64, bitwise inputs; sums to
    1e-12 relative, the device reduction order differs)."""
    from lbman3d_amd import simio
    print("case simio run loop", flush=True)
    dims = (12, 10, 8)
    cfg = abi.make_config(*dims, abi.ChannelConfig)
    q = abi.seed_director_field(5, *dims)
    path = tmp_path / "lbm.log"
    saved = []
    done = simio.run(cfg, q, 30, 10, "Channel", str(path), on_save=lambda t, s: saved.append(t))
    assert done == 30 and saved == [0, 10, 20]
    lines = [ln for ln in path.read_text().splitlines() if ln.startswith("Time ")]
    assert len(lines) == 3
    o = oracle_lib.Oracle(cfg)
    o.set_state(q=q)
    o.step(20)
    ls = o.log_sums()
    fields = dict(kv.split(": ") for kv in lines[2].split(": ", 1)[1].split(", "))
    assert float(fields["Mass"]) == pytest.approx(ls["mass"], rel=1e-12)
    assert float(fields["Kinetic Energy"]) == pytest.approx(ls["kinetic"], rel=1e-10)
    assert float(fields["Max |u|"]) == pytest.approx(ls["umax"], rel=1e-12)
    assert float(fields["Nematic Energy"]) == pytest.approx(o.free_energy(), rel=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_SHIFTED])
def test_device_nonfinite_flag_and_diverged_stop(tmp_path, precision):
    """A NaN in the simulated state reaches the device reductions: the
    nonfinite flag is set and the main loop writes the DIVERGED line and
    stops (src/sim_io.cc:133-165, src/main.cc:17-20). The NaN is data only:
    IEEE arithmetic carries it and nothing faults. max |u| skips a NaN, as the
    reference's `if (u2 > umax2)` does (src/sim_io.cc:96)."""
    from lbman3d_amd import simio
    from lbman3d_amd.solver import DeviceSolver
    print(f"case nonfinite precision={precision}", flush=True)
    dims = (12, 10, 8)
    V = dims[0] * dims[1] * dims[2]
    cfg = abi.make_config(*dims, abi.ChannelConfig, precision=precision)
    q = abi.seed_director_field(5, *dims)
    rho = np.ones(V)
    rho[37] = math.nan
    with DeviceSolver(cfg) as s:
        s.upload_state(rho=rho, q=q)
        d = s.diagnostics()
        assert d["nonfinite"] == 1 and math.isnan(d["mass"])
        assert math.isfinite(d["umax"])
        s.upload_state(rho=np.ones(V))        # populations Feq(1, u) again: finite
        d = s.diagnostics()
        assert d["nonfinite"] == 0 and d["mass"] == pytest.approx(V, rel=1e-6)

    # A NaN in one cell of Q spreads through the force into the populations;
    # the Log at step 10 sees a NaN mass.
    qbad = q.copy()
    qbad[2, 123] = math.nan
    path = tmp_path / "lbm.log"
    done = simio.run(cfg, qbad, 30, 10, "Channel", str(path))
    assert done == 10
    text = path.read_text().splitlines()
    assert text[-2] == "DIVERGED at time step 10 — aborting."
    assert text[-1] == "LBM program exited."
    assert sum(1 for ln in text if ln.startswith("Time ")) == 2
    line10 = [ln for ln in text if ln.startswith("Time 10:")][0]
    assert "Mass: nan," in line10 or "Mass: -nan," in line10
