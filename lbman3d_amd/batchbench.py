"""Batched sweep: 64 independent FP32 simulations in one launch per step.

Workload (drawn from a fixed seed): box sides uniform in {24, 32, 48, 64},
half FullyPeriodicConfig and half SlitConfig (free-slip walls in x), activity
zeta uniform in [0.5, 2] x 0.023, flow alignment lambda uniform in [0.3, 0.7],
centred advection, each simulation seeded with its own random director field
(S0 = 0.5 +- 0.01). The batch is stepped by the persistent batched kernel from
a replayed hipGraph; the same 64 simulations are then stepped one after
another as single handles (each on its own autotuned fused schedule), and
once more one after another on the FP64 plain path. The aggregate MLUPS of
all three are reported together with the number of simulations that
diverged (non-finite mass or momentum) in the batched run.

Accuracy leg: the batch is re-seeded and stepped again while every
simulation also runs on the FP64 plain path (the reference's arithmetic), on
the FP64 path from Qxx perturbed by one ulp, and on the FP64 path from Q
rounded to FP32 (how far the dynamics alone carries a difference of FP32
size, injected once). Every
`--accuracy-every` steps the relative L2 error of Q, u and rho - rho0 of each
batched simulation against its FP64 run is recorded (worst and median over
the simulations, and per simulation).

usage: python -m lbman3d_amd.batchbench [--steps 5000] [--sims 64] [--accuracy-every 1000]
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import abi
from .solver import BatchSolver, DeviceSolver, seed_director_field


def workload(nsims=64, seed=7):
    rng = np.random.default_rng(seed)
    sides = [24, 32, 48, 64]
    cfgs, seeds, meta = [], [], []
    for i in range(nsims):
        dims = tuple(int(x) for x in rng.choice(sides, 3))
        bc = abi.FullyPeriodicConfig if i % 2 == 0 else abi.SlitConfig
        zeta = float(rng.uniform(0.5, 2.0)) * 0.023
        lam = float(rng.uniform(0.3, 0.7))
        cfgs.append(abi.make_config(*dims, bc, abi.Params(ALPHA=zeta, LAMBDA=lam), advection=abi.ADV_CENTRED,
                                    precision=abi.FP32_SHIFTED))
        seeds.append(10_000 + i)
        meta.append({"dims": dims, "bc": bc.name, "zeta": zeta, "lambda": lam})
    return cfgs, seeds, meta


def rel_l2(a, b):
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))


def accuracy(cfgs, qs, steps, every):
    """Per-simulation error of the batched FP32 step against FP64 runs of the
    same configurations and seeds, every `every` steps up to `steps`."""
    n = len(cfgs)
    f64 = [DeviceSolver(abi.config_copy(c, precision=abi.FP64)) for c in cfgs]
    ulp = [DeviceSolver(abi.config_copy(c, precision=abi.FP64)) for c in cfgs]
    r32 = [DeviceSolver(abi.config_copy(c, precision=abi.FP64)) for c in cfgs]
    rows = []
    try:
        with BatchSolver(cfgs) as b:
            for i, q in enumerate(qs):
                b.upload_state(i, q=q)
                f64[i].upload_state(q=q)
                qp = q.copy()
                qp[0] = np.nextafter(qp[0], np.inf)
                ulp[i].upload_state(q=qp)
                r32[i].upload_state(q=q.astype(np.float32).astype(np.float64))
            done = 0
            while done < steps:
                k = min(every, steps - done)
                b.step(k, use_graph=True)
                for s in f64 + ulp + r32:
                    s.step(k)
                done += k
                per = []
                for i in range(n):
                    ref, st, tw = f64[i].state(), b.state(i), ulp[i].state()
                    e = lambda x: {"q": rel_l2(x["q"], ref["q"]), "u": rel_l2(x["u"], ref["u"]),
                                   "rho": rel_l2(x["rho"] - 1.0, ref["rho"] - 1.0)}
                    per.append({"fp32_batch": e(st), "fp64_1ulp": e(tw), "fp64_from_fp32_q": e(r32[i].state()),
                                "fp64_nonfinite": f64[i].diagnostics()["nonfinite"]})
                row = {"step": done}
                for leg in ("fp32_batch", "fp64_1ulp", "fp64_from_fp32_q"):
                    for f in ("q", "u", "rho"):
                        v = np.array([p[leg][f] for p in per])
                        row[f"{leg}_{f}_max"] = float(np.nanmax(v)) if np.isfinite(v).any() else None
                        row[f"{leg}_{f}_median"] = float(np.nanmedian(v)) if np.isfinite(v).any() else None
                row["fp32_batch_q_per_sim"] = [p["fp32_batch"]["q"] for p in per]
                row["sims_within_1e-5"] = int(sum(all(x < 1e-5 for x in p["fp32_batch"].values()) for p in per))
                row["fp64_diverged"] = [i for i, p in enumerate(per) if p["fp64_nonfinite"]]
                print(json.dumps({k: v for k, v in row.items() if k != "fp32_batch_q_per_sim"}), flush=True)
                rows.append(row)
    finally:
        for s in f64 + ulp + r32:
            s.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--sims", type=int, default=64)
    ap.add_argument("--seq-steps", type=int, default=200, help="steps per simulation in the sequential leg")
    ap.add_argument("--accuracy-every", type=int, default=1000, help="0 skips the accuracy leg")
    a = ap.parse_args()
    cfgs, seeds, meta = workload(a.sims)
    qs = [seed_director_field(sd, c.nx, c.ny, c.nz) for sd, c in zip(seeds, cfgs)]
    cells = sum(c.nx * c.ny * c.nz for c in cfgs)

    with BatchSolver(cfgs) as b:
        for i, q in enumerate(qs):
            b.upload_state(i, q=q)
        tune = b.autotune(5)
        b.step(2)
        b.time_steps(4)                          # graph capture outside the timed region
        t0 = time.perf_counter()
        ms = b.time_steps(a.steps)
        wall = time.perf_counter() - t0
        diags = [b.diagnostics(i) for i in range(len(cfgs))]
    diverged = [i for i, d in enumerate(diags) if d["nonfinite"] or not np.isfinite(d["umax"])]
    batch_mlups = cells * a.steps / (ms * 1e-3) / 1e6

    seq_ms = 0.0
    for c, q in zip(cfgs, qs):
        with DeviceSolver(c) as s:
            s.upload_state(q=q)
            s.autotune(3)
            s.step(2)
            seq_ms += s.time_steps(a.seq_steps)
    seq_mlups = cells * a.seq_steps / (seq_ms * 1e-3) / 1e6

    # the FP64 plain path (reference-parity kernels), same simulations one by one
    f64_ms = 0.0
    for c, q in zip(cfgs, qs):
        with DeviceSolver(abi.config_copy(c, precision=abi.FP64)) as s:
            s.upload_state(q=q)
            s.step(2)
            f64_ms += s.time_steps(a.seq_steps)
    f64_mlups = cells * a.seq_steps / (f64_ms * 1e-3) / 1e6
    acc = accuracy(cfgs, qs, a.steps, a.accuracy_every) if a.accuracy_every > 0 else None
    print(json.dumps({"workload": "batched sweep", "sims": len(cfgs), "cells": cells, "steps": a.steps,
                      "batch_ms_per_step": ms / a.steps, "batch_mlups": batch_mlups, "batch_wall_s": wall,
                      "batch_tile_ms": tune, "sequential_ms_per_step": seq_ms / a.seq_steps,
                      "sequential_mlups": seq_mlups, "speedup": batch_mlups / seq_mlups,
                      "sequential_fp64_ms_per_step": f64_ms / a.seq_steps, "sequential_fp64_mlups": f64_mlups,
                      "speedup_vs_fp64_plain": batch_mlups / f64_mlups,
                      "diverged": diverged, "umax": [d["umax"] for d in diags], "accuracy": acc,
                      "sims_meta": meta}), flush=True)


if __name__ == "__main__":
    main()
