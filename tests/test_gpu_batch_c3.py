"""The batched sweep's own workload against the FP64 oracle (needs a GPU).

The 64 simulations of the batched sweep (box sides from {24, 32, 48, 64},
even ones FullyPeriodicConfig, odd ones SlitConfig with free-slip walls in x
(src/sim_config.h:19-27), zeta uniform in [0.5, 2] x 0.023, lambda uniform in
[0.3, 0.7], each with its own seeded director field) are advanced together
by the batched kernel on the autotuned tile, replayed from a hipGraph. Every
simulation is compared with the oracle run of the same configuration and
seed in FP64 (the reference's arithmetic, one configuration per run as the
reference builds it, src/device_solver.cu:123-125), not with other device
runs. The oracle simulations run concurrently on the host's cores.

Bar: relative L2 error of Q, u and rho - rho0 below 1e-5 for every
simulation after 100 steps (before the active flows amplify rounding).
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

from lbman3d_amd import abi

pytestmark = pytest.mark.gpu

STEPS = 100
WORKERS = max(1, min(16, len(os.sched_getaffinity(0))))


def rel_l2(a, b):
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))


def test_c3_batch_against_oracle(oracle_lib):
    from lbman3d_amd.batchbench import workload
    from lbman3d_amd.solver import BatchSolver, batch_tile_name
    cfgs, seeds, meta = workload(64)
    qs = [abi.seed_director_field(sd, c.nx, c.ny, c.nz) for sd, c in zip(seeds, cfgs)]
    assert sum(m["bc"] == "SlitFreeSlip" for m in meta) == 32
    print(f"case C3 batch: 64 simulations, {STEPS} steps", flush=True)
    with BatchSolver(cfgs) as b:
        for i, q in enumerate(qs):
            b.upload_state(i, q=q)
        b.autotune(3)
        tile = batch_tile_name(b.get_tile())
        b.step(STEPS, use_graph=True)
        got = [b.state(i) for i in range(len(cfgs))]
    print(f"C3 batch on tile {tile}; oracle on {WORKERS} workers", flush=True)

    def oracle(i):
        o = oracle_lib.Oracle(abi.config_copy(cfgs[i], precision=abi.FP64), threads=1)
        o.set_state(q=qs[i])
        o.step(STEPS)
        return o.state()

    with cf.ThreadPoolExecutor(WORKERS) as ex:
        refs = list(ex.map(oracle, range(len(cfgs))))
    worst = {"q": 0.0, "u": 0.0, "rho": 0.0}
    bad = []
    for i, (st, ref) in enumerate(zip(got, refs)):
        e = {"q": rel_l2(st["q"], ref["q"]), "u": rel_l2(st["u"], ref["u"]),
             "rho": rel_l2(st["rho"] - 1.0, ref["rho"] - 1.0)}
        for k in worst:
            worst[k] = max(worst[k], e[k])
        if not all(v < 1e-5 for v in e.values()):
            bad.append((i, meta[i], e))
    print(f"C3 worst relative L2 over 64 simulations after {STEPS} steps: {worst}", flush=True)
    assert not bad, bad[:5]
