"""Throughput of the lbman3d step on one MI355X.

Default workload (the bandwidth run): 256^3 FullyPeriodicConfig, FP32 with
shifted populations, upwind advection, zeta = 0.046 (2x the params.h
default), lambda = 0.7, seeded random director with S0 = 0.5 +- 0.01, rho = 1,
u = 0. A "step" is one full coupled step (Q update + body force + D3Q15
collide-stream) over the whole grid; inputs are resident in HBM before the
timed region starts.

Prints ONE JSON line: MLUPS (whole job), ms per step, the roofline of the
dominant kernel (algorithmic bytes per launch / mean launch duration measured
with HIP events on the solver's stream, against the 8 TB/s HBM peak), and a
CPU baseline: the reference's own OpenMP CPU solver (oracle/_ref/ref_c2,
compiled from its sources) timed on 10 warm-up + 30 steps of the same workload
on the host cores this process may use (CPU model and core counts stated).
The FP32 result's relative L2 error against the FP64 step (Q, u, rho - rho0)
at 250 and 1000 steps is reported beside it.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-accuracy]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per lattice update: populations 15 read + 15 write,
# Q 5 + 5, u 3 + 3 = 46 values (FP32: 184 B, FP64: 368 B). For the kernels of
# the two-kernel split, the compulsory traffic of each kernel on its own.
ALGO_BYTES = {
    "fused_step_f32": 46 * 4,
    "fused_step_f64": 46 * 8,
    "qstep_f32": (8 + 16) * 4,          # Q, u in; Q', Sigma, Tau, F out
    "force_lbm_f32": (5 + 8 + 3 + 15 + 15 + 3) * 4,
    "qstep_f64": (8 + 16) * 8,
    "force_lbm_f64": (5 + 8 + 3 + 15 + 15 + 3 + 1 + 3) * 8,
}
STEP_BYTES = {"fp32": 46 * 4, "fp64": 46 * 8}
PROFILE_DIR = os.path.join(ROOT, "profiles")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_cfg(n, precision, upwind=True):
    from lbman3d_amd import abi
    return abi.make_config(n, n, n, abi.FullyPeriodicConfig, abi.Params(ALPHA=0.046, LAMBDA=0.7),
                           advection=abi.ADV_UPWIND if upwind else abi.ADV_CENTRED, precision=precision)


def time_solver(cfg, q, steps, warmup, schedule=None, autotune=False):
    from lbman3d_amd.solver import DeviceSolver
    s = DeviceSolver(cfg)
    s.upload_state(q=q)
    tuned = None
    if autotune:
        tuned = s.autotune(5)
    if schedule is not None:
        s.set_schedule(schedule)
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    ev_ms = s.time_steps(steps)
    wall = time.perf_counter() - t0
    prof = s.profile_kernels(min(20, steps))
    diag = s.diagnostics()
    return s, ev_ms, wall, prof, diag, tuned


def roofline(prof, V, schedule_name=None, step_ms=None):
    """Roofline of the dominant kernel. traffic is the PMC-measured HBM bytes
    per launch of the schedule that ran, from the committed profile (null
    when that schedule has no measurement).

    The kernel's average launch duration: when the step is that one launch
    (prof names one kernel) and step_ms is given, the timed region's
    HIP-event time divided by its launches, step_ms; otherwise the mean of
    event pairs around single launches, which also hold each launch's
    dispatch (kept beside it as kernel_ms_single_launch_events)."""
    name = max(prof, key=lambda k: prof[k])
    ms = prof[name]
    source = "event pairs around single launches"
    if step_ms is not None and len(prof) == 1:
        ms = step_ms
        source = "timed-region HIP events / launches"
    algo = ALGO_BYTES.get(name)
    if algo is None:
        return name, None
    achieved = algo * V / (ms * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(PROFILE_DIR, f"{name}_256cube_traffic.json")
    if os.path.exists(pmc):
        with open(pmc) as fp:
            entry = json.load(fp).get("by_schedule", {}).get(schedule_name)
        if entry:
            traffic = entry["hbm_bytes_per_launch"]
    return name, {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                  "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": name,
                  "kernel_ms": round(ms, 4), "kernel_ms_source": source,
                  "kernel_ms_single_launch_events": round(prof[name], 4), "algorithmic_bytes_per_launch": algo * V}


def host_cpu():
    """CPU model and physical core count of this host (lscpu), and the
    threads this process may use (affinity, capped by OMP_NUM_THREADS)."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, env=dict(os.environ, LC_ALL="C"),
                             timeout=30).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        if kv.get("Core(s) per socket") and kv.get("Socket(s)"):
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
        info["logical_cpus"] = int(kv["CPU(s)"]) if kv.get("CPU(s)") else None
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    allowed = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    info["threads"] = min(allowed, int(env)) if env.isdigit() and int(env) > 0 else allowed
    info["allowed_cpus"] = allowed
    return info


def cpu_baseline(q, n, steps=30, warmup=10):
    """The reference's OpenMP CPU solver (oracle/_ref/ref_c2: 256^3
    FullyPeriodic, upwind, zeta 0.046, lambda 0.7, compiled from its sources
    with the thread count read at run time) on the same workload, under the
    protocol of src/benchmark.cc:9-50: `warmup` untimed steps, then `steps`
    steps timed one by one with the QTensor / LBM split. Threads are bound
    close, one per core (src/mpi/CLAUDE.md "Baseline invocation"), as many as
    this process may use."""
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_c2")
    if not os.path.exists(exe) or n != 256:
        return None
    cpu = host_cpu()
    with tempfile.TemporaryDirectory(prefix="lbmcpu_") as td:
        fin = os.path.join(td, "in.bin")
        with open(fin, "wb") as fp:
            np.array([n, n, n, 0, 0], dtype=np.int32).tofile(fp)
            np.ascontiguousarray(q, dtype=np.float64).reshape(-1).tofile(fp)
        env = dict(os.environ, OMP_NUM_THREADS=str(cpu["threads"]), OMP_PROC_BIND="close", OMP_PLACES="cores")
        log(f"cpu baseline: reference solver, {cpu['threads']} threads, {warmup} + {steps} steps")
        r = subprocess.run([exe, "FullyPeriodic", "full", "0", str(steps), fin, "-", str(warmup)],
                           capture_output=True, text=True, env=env, timeout=600)
    m = re.search(r"ref: (\d+) steps of (\d+) cells in ([0-9.]+) ms \(([0-9.]+) MLUPS, (\d+) threads; "
                  r"(\d+) warm-up steps; QTensor ([0-9.]+) ms/step, LBM ([0-9.]+) ms/step\)", r.stderr)
    if r.returncode != 0 or not m:
        log(f"cpu baseline failed: {r.stderr[-400:]}")
        return None
    return {"value": float(m.group(4)), "unit": "MLUPS", "cores": int(m.group(5)), "kind": "reference",
            "sample": f"{m.group(6)} warm-up + {m.group(1)} timed steps of the 256^3 FP64 workload "
                      f"(reference CPU solver, {m.group(5)} OpenMP threads bound close to cores, "
                      f"{float(m.group(3)):.0f} ms timed)",
            "qtensor_ms_per_step": float(m.group(7)), "lbm_ms_per_step": float(m.group(8)),
            "cpu_model": cpu.get("model"), "host_physical_cores": cpu.get("physical_cores"),
            "host_logical_cpus": cpu.get("logical_cpus"), "cpus_allotted": cpu["allowed_cpus"]}


def accuracy(n, q, checkpoints=(250, 1000)):
    """Relative L2 error of the FP32 step against the FP64 step (the
    reference's arithmetic bit for bit) from the same seeded state, in Q, u
    and rho - rho0, at the given step counts. Beside it, how far the
    dynamics alone carries small differences: the FP64 step started from Qxx
    perturbed by one ulp in every cell, and the FP64 step started from Q
    rounded to FP32 (a difference of FP32 size, injected once)."""
    from lbman3d_amd import abi
    from lbman3d_amd.solver import DeviceSolver
    qp = q.copy()
    qp[0] = np.nextafter(qp[0], np.inf)
    sims = {"fp64": DeviceSolver(make_cfg(n, abi.FP64)), "fp32": DeviceSolver(make_cfg(n, abi.FP32_SHIFTED)),
            "fp64_1ulp": DeviceSolver(make_cfg(n, abi.FP64)), "fp64_from_fp32_q": DeviceSolver(make_cfg(n, abi.FP64))}
    sims["fp64"].upload_state(q=q)
    sims["fp32"].upload_state(q=q)
    sims["fp64_1ulp"].upload_state(q=qp)
    sims["fp64_from_fp32_q"].upload_state(q=q.astype(np.float32).astype(np.float64))
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out, done = {}, 0
    try:
        for cp in checkpoints:
            for s in sims.values():
                s.step(cp - done)
            done = cp
            ref = sims["fp64"].state()
            row = {}
            for k in ("fp32", "fp64_1ulp", "fp64_from_fp32_q"):
                st = sims[k].state()
                row[k] = {"q": rel(st["q"], ref["q"]), "u": rel(st["u"], ref["u"]),
                          "rho_minus_rho0": rel(st["rho"] - 1.0, ref["rho"] - 1.0)}
            row["fp64_umax"] = sims["fp64"].diagnostics()["umax"]
            out[str(cp)] = row
            log(f"accuracy at step {cp}: {row}")
    finally:
        for s in sims.values():
            s.close()
    return {"vs": "FP64 step from the same seeded state (bitwise equal to the reference CPU solver)",
            "by_step": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256, help="grid edge")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--schedule", type=int, default=None, help="force an FP32 kernel schedule")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (requested --gpus {args.gpus})",
              file=sys.stderr)
        sys.exit(2)

    from lbman3d_amd import abi
    from lbman3d_amd.solver import seed_director_field, tile_count, device_info
    n = args.n
    V = n * n * n
    log(device_info())
    log(f"seeding {n}^3 director field")
    q = seed_director_field(2024, n, n, n, 0.5, 0.01)

    autotune = args.schedule is None and tile_count() > 0
    log("timing FP32 shifted step")
    s, ev_ms, wall, prof, diag, tuned = time_solver(make_cfg(n, abi.FP32_SHIFTED), q, args.steps, args.warmup,
                                                    schedule=args.schedule, autotune=autotune)
    schedule = s.get_schedule()
    if schedule == 0 and tuned:
        # autotune order is schedule 1, 2, ...; "auto" runs the fastest
        names = list(tuned)
        best = min(names, key=tuned.get)
        schedule_used = {"schedule": names.index(best) + 1, "name": best}
    else:
        from lbman3d_amd.solver import schedule_name
        schedule_used = {"schedule": schedule, "name": schedule_name(schedule)}
    s.close()
    ms_step = ev_ms / args.steps
    mlups = V / (ms_step * 1e-3) / 1e6
    kname, roof = roofline(prof, V, schedule_used["name"], step_ms=ms_step)
    step_frac = STEP_BYTES["fp32"] * V / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS
    log(f"FP32: {ms_step:.4f} ms/step, {mlups:.1f} MLUPS, kernels {prof}")

    from lbman3d_amd.solver import bandwidth_probe
    copy_gbs, copy_ms = bandwidth_probe(1 << 30, 20)
    log(f"copy kernel: {copy_gbs:.0f} GB/s")
    secondary = {"fp32_step_hbm_frac": round(step_frac, 4), "fp32_kernels_ms": prof,
                 "hbm_copy_GBs": round(copy_gbs, 1),
                 "fp32_frac_of_copy_bw": round(STEP_BYTES["fp32"] * V / (ms_step * 1e-3) / 1e9 / copy_gbs, 4),
                 "fp32_wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
                 "fp32_umax": diag["umax"], "fp32_nonfinite": diag["nonfinite"]}
    if tuned:
        secondary["autotune_ms_per_step"] = tuned
    if not args.no_fp64:
        log("timing FP64 plain step")
        k64 = max(10, args.steps // 5)
        s64, ev64, _, prof64, _, _ = time_solver(make_cfg(n, abi.FP64), q, k64, 5)
        name64 = s64.schedule_name(0)
        # every FP64 schedule: the two-kernel split and the fused tiles
        by_sched = {}
        sched = 1
        while True:
            try:
                nm = s64.schedule_name(sched)
            except Exception:
                break
            s64.set_schedule(sched)
            s64.step(3)
            by_sched[nm] = round(s64.time_steps(k64) / k64, 4)
            sched += 1
        s64.close()
        ms64 = ev64 / k64
        k64name, roof64 = roofline(prof64, V, name64.replace("auto:", "", 1), step_ms=ms64)
        secondary.update({"fp64_ms_per_step": round(ms64, 4), "fp64_mlups": round(V / (ms64 * 1e-3) / 1e6, 1),
                          "fp64_schedule": name64,
                          "fp64_step_hbm_frac": round(STEP_BYTES["fp64"] * V / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "fp64_kernels_ms": prof64, "fp64_roofline": roof64, "fp64_schedules_ms_per_step": by_sched})

    if not args.no_accuracy:
        log("accuracy: FP32 and 1-ulp FP64 against FP64")
        secondary["accuracy"] = accuracy(n, q)
    cpu = None if args.no_cpu_baseline else cpu_baseline(q, n)
    out = {
        "metric": "MLUPS and % peak HBM BW; FP32/batched/closure vs FP64 plain at equal error",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded random director field, rho=1, u=0",
        "config": {"workload": f"{n}^3 FullyPeriodicConfig coupled LB+Q-tensor step, FP32 shifted populations, "
                               "upwind advection, zeta=0.046, lambda=0.7, S0=0.5+-0.01",
                   "grid": [n, n, n], "schedule": schedule, "schedule_used": schedule_used,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
