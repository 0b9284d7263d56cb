"""Accuracy of the FP32 shifted-population step against the FP64 step.

The FP64 step is bit-identical to the reference CPU solver (tests/), so it is
the yardstick. Three runs start from the same seeded state:
  fp64      the FP64 plain step,
  fp32      the FP32 shifted-population step (auto schedule),
  fp64-ulp  the FP64 step from the same state with Qxx perturbed by one ulp
            in every cell: the chaos floor, i.e. how far the dynamics itself
            carries a rounding-level difference after n steps.
Every `every` steps the relative L2 error of Q, u and rho - rho0 of fp32 and
fp64-ulp against fp64 is printed as one JSON line.

usage: python -m lbman3d_amd.accuracy --workload channel          (20x100x100, 10^4 steps)
       python -m lbman3d_amd.accuracy --workload channel-passive  (the same with zeta = 0)
       python -m lbman3d_amd.accuracy --workload bandwidth        (256^3, 2000 steps)
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import abi
from .solver import DeviceSolver, seed_director_field

WORKLOADS = {
    # ChannelConfig 20x100x100, params.h defaults with lambda 0.5, centred.
    "channel": dict(dims=(20, 100, 100), bc=abi.ChannelConfig, params=abi.Params(LAMBDA=0.5),
                    adv=abi.ADV_CENTRED, steps=10000, every=1000, seed=1),
    # The same channel without activity (zeta 0): relaxational, not chaotic,
    # so the FP32 error is the arithmetic's own and is not amplified.
    "channel-passive": dict(dims=(20, 100, 100), bc=abi.ChannelConfig, params=abi.Params(LAMBDA=0.5, ALPHA=0.0),
                            adv=abi.ADV_CENTRED, steps=10000, every=1000, seed=1),
    # FullyPeriodic 256^3, zeta 2x default, lambda 0.7, upwind.
    "bandwidth": dict(dims=(256, 256, 256), bc=abi.FullyPeriodicConfig,
                      params=abi.Params(ALPHA=0.046, LAMBDA=0.7), adv=abi.ADV_UPWIND, steps=2000, every=250,
                      seed=2024),
}


def rel_l2(a, b):
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))


def run(workload: str, steps: int | None = None, every: int | None = None):
    w = WORKLOADS[workload]
    dims = w["dims"]
    steps = steps or w["steps"]
    every = every or w["every"]
    q0 = seed_director_field(w["seed"], *dims, 0.5, 0.01)
    qp = q0.copy()
    qp[0] = np.nextafter(qp[0], np.inf)
    mk = lambda prec: abi.make_config(*dims, w["bc"], w["params"], advection=w["adv"], precision=prec)
    sims = {"fp64": DeviceSolver(mk(abi.FP64)), "fp32": DeviceSolver(mk(abi.FP32_SHIFTED)),
            "fp64-ulp": DeviceSolver(mk(abi.FP64))}
    sims["fp64"].upload_state(q=q0)
    sims["fp32"].upload_state(q=q0)
    sims["fp64-ulp"].upload_state(q=qp)
    rho0 = w["params"].kDensity
    ms = {k: 0.0 for k in sims}
    done = 0
    while done < steps:
        n = min(every, steps - done)
        for k, s in sims.items():
            ms[k] += s.time_steps(n)
        done += n
        ref = sims["fp64"].state()
        line = {"workload": workload, "step": done}
        for k in ("fp32", "fp64-ulp"):
            st = sims[k].state()
            line[k] = {"Q": rel_l2(st["q"], ref["q"]), "u": rel_l2(st["u"], ref["u"]),
                       "rho": rel_l2(st["rho"] - rho0, ref["rho"] - rho0)}
        d = sims["fp64"].diagnostics()
        line["fp64_umax"] = d["umax"]
        line["fp64_nonfinite"] = d["nonfinite"]
        print(json.dumps(line), flush=True)
    V = int(np.prod(dims))
    summary = {"workload": workload, "steps": steps,
               "ms_per_step": {k: v / steps for k, v in ms.items()},
               "mlups": {k: V * steps / (v * 1e-3) / 1e6 for k, v in ms.items()}}
    print(json.dumps(summary), flush=True)
    for s in sims.values():
        s.close()
    return summary


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="channel")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--every", type=int, default=None)
    a = ap.parse_args()
    t0 = time.time()
    run(a.workload, a.steps, a.every)
    print(json.dumps({"wall_s": time.time() - t0}), flush=True)


if __name__ == "__main__":
    main()
