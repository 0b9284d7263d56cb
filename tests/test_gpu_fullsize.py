"""Parity at the full sizes of the channel and bandwidth workloads (needs a GPU).

- Channel 20 x 100 x 100 (no-slip/Neumann walls in x and y, periodic z),
  lambda 0.5, centred advection: the reference's own stability workload
  (src/sim_config.h:60-66). FP64 must equal the oracle bit for bit after 200
  steps; FP32 (automatic and forced fused schedules) must stay within 1e-5
  relative L2 in Q, u and rho - rho0 of the oracle after 1000 steps, before
  the flow turns chaotic.
- FullyPeriodic 256^3, zeta 0.046, lambda 0.7, upwind: the bandwidth
  workload. FP64 must equal the oracle bit for bit after 5 steps. The FP32
  step on the schedules the bandwidth run uses (p2, p2z and their forms with
  unconditional memory instructions, 64 x 4 x 128 tiles, two z chunks per
  column, so the chunk boundary at z = 128 is crossed) must stay within 1e-6 of the oracle after those 5 steps, and
  within 1e-5 of the FP64 step after 250 steps (the FP64 step at this size
  being the oracle's bits, by the first check).

The oracle runs on the host's cores (up to 16 OpenMP threads).
"""
import os

import numpy as np
import pytest

from lbman3d_amd import abi

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
C1_DIMS = (20, 100, 100)
C2_N = 256


def rel_l2(a, b):
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))


def errors(st, ref, rho0=1.0):
    return {"q": rel_l2(st["q"], ref["q"]), "u": rel_l2(st["u"], ref["u"]),
            "rho": rel_l2(st["rho"] - rho0, ref["rho"] - rho0)}


def schedule(name):
    from lbman3d_amd.solver import schedule_name, tile_count
    for s in range(2, 2 + tile_count()):
        if schedule_name(s) == "fused:" + name:
            return s
    raise KeyError(name)


def c1_config(precision):
    return abi.make_config(*C1_DIMS, abi.ChannelConfig, abi.Params(LAMBDA=0.5), advection=abi.ADV_CENTRED,
                           precision=precision)


def c2_config(precision):
    return abi.make_config(C2_N, C2_N, C2_N, abi.FullyPeriodicConfig, abi.Params(ALPHA=0.046, LAMBDA=0.7),
                           advection=abi.ADV_UPWIND, precision=precision)


def device_run(cfg, q, steps, sched=None):
    from lbman3d_amd.solver import DeviceSolver
    with DeviceSolver(cfg) as s:
        if sched is not None:
            s.set_schedule(sched)
        s.upload_state(q=q)
        s.step(steps)
        return s.state()


def oracle_run(oracle_lib, cfg, q, steps):
    o = oracle_lib.Oracle(cfg, threads=THREADS)
    o.set_state(q=q)
    o.step(steps)
    return o.state()


@pytest.fixture(scope="module")
def c1_q():
    return abi.seed_director_field(1, *C1_DIMS)


@pytest.fixture(scope="module")
def c1_oracle_1000(oracle_lib, c1_q):
    print("case C1 oracle 1000 steps", flush=True)
    return oracle_run(oracle_lib, c1_config(abi.FP64), c1_q, 1000)


def test_c1_fp64_bitwise_200_steps(oracle_lib, c1_q):
    print("case C1 FP64 200 steps", flush=True)
    ref = oracle_run(oracle_lib, c1_config(abi.FP64), c1_q, 200)
    st = device_run(c1_config(abi.FP64), c1_q, 200)
    for k in ("rho", "u", "q"):
        np.testing.assert_array_equal(st[k], ref[k], err_msg=k)


@pytest.mark.parametrize("sched", [None, "two-kernel", "p2:64x4x16", "p2z:64x4x32", "pf-nt:32x8x16"])
def test_c1_fp32_within_1e5_after_1000_steps(c1_oracle_1000, c1_q, sched):
    print(f"case C1 FP32 1000 steps schedule={sched}", flush=True)
    s = None if sched is None else (1 if sched == "two-kernel" else schedule(sched))
    st = device_run(c1_config(abi.FP32_SHIFTED), c1_q, 1000, s)
    e = errors(st, c1_oracle_1000)
    print(f"C1 FP32 {sched}: {e}", flush=True)
    assert e["q"] < 1e-5 and e["u"] < 1e-5 and e["rho"] < 1e-5, e


@pytest.fixture(scope="module")
def c2_q():
    return abi.seed_director_field(2024, C2_N, C2_N, C2_N, 0.5, 0.01)


@pytest.fixture(scope="module")
def c2_oracle_5(oracle_lib, c2_q):
    print("case C2 oracle 5 steps", flush=True)
    return oracle_run(oracle_lib, c2_config(abi.FP64), c2_q, 5)


def test_c2_fp64_bitwise_5_steps(c2_oracle_5, c2_q):
    print("case C2 FP64 5 steps", flush=True)
    st = device_run(c2_config(abi.FP64), c2_q, 5)
    for k in ("rho", "u", "q"):
        np.testing.assert_array_equal(st[k], c2_oracle_5[k], err_msg=k)


@pytest.mark.parametrize("sched", [None, "p2:64x4x128", "p2z:64x4x128", "p2u:64x4x128", "p2uz:64x4x128"])
def test_c2_fp32_bench_schedules_against_oracle(c2_oracle_5, c2_q, sched):
    print(f"case C2 FP32 5 steps schedule={sched}", flush=True)
    st = device_run(c2_config(abi.FP32_SHIFTED), c2_q, 5, None if sched is None else schedule(sched))
    e = errors(st, c2_oracle_5)
    print(f"C2 FP32 5 steps {sched}: {e}", flush=True)
    assert e["q"] < 1e-6 and e["u"] < 1e-5 and e["rho"] < 1e-5, e


@pytest.mark.parametrize("sched", ["p2:64x4x128", "p2z:64x4x128", "p2uz:64x4x128"])
def test_c2_fp32_against_fp64_after_250_steps(c2_q, sched):
    print(f"case C2 FP32 vs FP64 250 steps schedule={sched}", flush=True)
    ref = device_run(c2_config(abi.FP64), c2_q, 250)
    st = device_run(c2_config(abi.FP32_SHIFTED), c2_q, 250, schedule(sched))
    e = errors(st, ref)
    print(f"C2 FP32 250 steps {sched}: {e}", flush=True)
    assert e["q"] < 1e-5 and e["u"] < 1e-5 and e["rho"] < 1e-5, e
