"""Learned closure for a 2x-coarsened active-nematic simulation.

Workload: a 128^3 FullyPeriodic FP64 run (the "fine" model) is box-filtered
2x2x2 onto 64^3; a 64^3 FP32 run (the "coarse" model) advances the filtered
state by one step, which stands for two fine steps. The coarse constants
follow acoustic scaling of the lattice units (dx and dt both doubled, so
lattice velocities are unchanged):

    tau_c - 1/2 = (tau - 1/2) / 2     (same kinematic viscosity)
    L_c = L / 4                       (L carries stress x length^2)
    Gamma_c = 2 Gamma                 (Gamma carries 1 / (stress x time))
    A, B, C, zeta, lambda unchanged   (stress units or dimensionless)

A network ``ClosureNet`` (lbm_closure_* in the C ABI) maps the coarse state
after a step (Q, u) to the difference between the filtered fine Q two fine
steps later and the coarse Q, and is added to Q after every coarse step.
Training pairs are teacher-forced: each coarse step starts from the filtered
fine state. The report gives the held-out one-step error with and without
the closure and a rollout from a held-out fine trajectory.

usage: python -m lbman3d_amd.closure [--iters 1500] [--samples 48]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import time

import numpy as np

from . import abi
from .solver import DeviceSolver, check, lib, seed_director_field

FLOP_PER_CELL = 2 * 27 * (8 * 32 + 32 * 32 + 32 * 5)   # one forward pass, per coarse cell


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class ClosureNet:
    """Host handle of the device network (lbm_closure_*)."""

    def __init__(self, nx, ny, nz, hidden=32, seed=1):
        self.shape = (nz, ny, nx)
        self.V = nx * ny * nz
        h = C.c_void_p()
        check(lib().lbm_closure_create(nx, ny, nz, hidden, seed, C.byref(h)))
        self.h = h
        n = C.c_int64()
        check(lib().lbm_closure_param_count(self.h, C.byref(n)))
        self.nparams = n.value

    def close(self):
        if getattr(self, "h", None):
            lib().lbm_closure_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def params(self):
        w = np.zeros(self.nparams)
        check(lib().lbm_closure_get_params(self.h, _p(w)))
        return w

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        assert w.size == self.nparams
        check(lib().lbm_closure_set_params(self.h, _p(w)))

    def set_normalization(self, in_scale, out_scale):
        i = np.ascontiguousarray(in_scale, dtype=np.float64)
        o = np.ascontiguousarray(out_scale, dtype=np.float64)
        check(lib().lbm_closure_set_normalization(self.h, _p(i), _p(o)))

    def add_sample(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(8, self.V)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(5, self.V)
        idx = C.c_int32()
        check(lib().lbm_closure_add_sample(self.h, _p(x), _p(y), C.byref(idx)))
        return idx.value

    def clear_samples(self):
        check(lib().lbm_closure_clear_samples(self.h))

    def train(self, iters, lr, seed=0):
        loss = C.c_double()
        check(lib().lbm_closure_train(self.h, iters, lr, seed, C.byref(loss)))
        return loss.value

    def loss_and_grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(8, self.V)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(5, self.V)
        loss = C.c_double()
        g = np.zeros(self.nparams)
        check(lib().lbm_closure_loss_and_grad(self.h, _p(x), _p(y), C.byref(loss), _p(g)))
        return loss.value, g

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(8, self.V)
        y = np.zeros((5, self.V))
        check(lib().lbm_closure_predict(self.h, _p(x), _p(y)))
        return y

    def apply(self, sim: DeviceSolver, alpha=1.0):
        check(lib().lbm_closure_apply(self.h, sim.h, alpha))

    def time_forward(self, n=20):
        ms = C.c_double()
        check(lib().lbm_closure_time_forward(self.h, n, C.byref(ms)))
        return ms.value


# ---- fine / coarse models -------------------------------------------------

def coarse_params(p: abi.Params) -> abi.Params:
    """Constants of the 2x-coarsened lattice under acoustic scaling."""
    return abi.Params(DT=p.DT, kDensity=p.kDensity, TAUF=0.5 + (p.TAUF - 0.5) / 2, L=p.L / 4, A=p.A, B=p.B,
                      C=p.C, LAMBDA=p.LAMBDA, GAMMA=2 * p.GAMMA, ALPHA=p.ALPHA, MU=p.MU)


def fine_config(n=128, params=None):
    return abi.make_config(n, n, n, abi.FullyPeriodicConfig, params or abi.Params(), advection=abi.ADV_UPWIND,
                           precision=abi.FP64)


def coarse_config(n=64, params=None):
    return abi.make_config(n, n, n, abi.FullyPeriodicConfig, coarse_params(params or abi.Params()),
                           advection=abi.ADV_UPWIND, precision=abi.FP32_SHIFTED)


def filter2(a, n):
    """2x2x2 box filter of channel-major fields (..., n^3) -> (..., (n/2)^3)."""
    lead = a.shape[:-1]
    g = a.reshape(*lead, n // 2, 2, n // 2, 2, n // 2, 2)
    k = len(lead)
    return g.mean(axis=(k + 1, k + 3, k + 5)).reshape(*lead, (n // 2) ** 3)


def filtered_state(sim: DeviceSolver, n):
    st = sim.state()
    return {"rho": filter2(st["rho"], n), "u": filter2(st["u"], n), "q": filter2(st["q"], n)}


def coarse_step(coarse: DeviceSolver, st):
    """One coarse step from a filtered state; returns the coarse (q, u)."""
    coarse.upload_state(rho=st["rho"], u=st["u"], q=st["q"])
    coarse.step(1)
    out = coarse.state()
    return out["q"], out["u"]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def generate_pairs(fine: DeviceSolver, coarse: DeviceSolver, n, count, spacing):
    """Teacher-forced pairs: x = coarse (q, u) after one step from the filtered
    fine state, y = filtered fine q two fine steps later minus coarse q."""
    xs, ys, base = [], [], []
    for _ in range(count):
        st = filtered_state(fine, n)
        fine.step(2)
        target = filter2(fine.state()["q"], n)
        qc, uc = coarse_step(coarse, st)
        xs.append(np.concatenate([qc, uc]))
        ys.append(target - qc)
        base.append((target, qc, st["q"]))
        fine.step(max(spacing - 2, 0))
    return xs, ys, base


def fine_targets(fine: DeviceSolver, n, steps, every):
    """Filtered fine start state and filtered fine Q every `every` coarse steps."""
    st = filtered_state(fine, n)
    targets = []
    for _ in range(steps // every):
        fine.step(2 * every)
        targets.append(filter2(fine.state()["q"], n))
    return st, targets


def rollout(coarse: DeviceSolver, net: ClosureNet | None, start, targets, every):
    """Coarse run from the filtered fine state (closure added after every
    step when `net` is given); relative L2 of Q against each target."""
    coarse.upload_state(rho=start["rho"], u=start["u"], q=start["q"])
    errs = []
    for t in targets:
        for _ in range(every):
            coarse.step(1)
            if net is not None:
                net.apply(coarse, 1.0)
        errs.append(rel_l2(coarse.state()["q"], t))
    return errs


def run(n_fine=128, samples=48, heldout=8, spacing=20, warmup=1500, iters=1500, lr=1e-3, roll_steps=40,
        roll_every=5, seed=11, log=print):
    nc = n_fine // 2
    t0 = time.perf_counter()
    p = abi.Params()
    report = {"workload": f"closure {nc}^3 coarse vs {n_fine}^3 fine", "fine": f"{n_fine}^3 FP64 upwind",
              "coarse": f"{nc}^3 FP32 upwind", "coarse_params": coarse_params(p).__dict__}
    with DeviceSolver(fine_config(n_fine, p)) as fine, DeviceSolver(coarse_config(nc, p)) as coarse:
        coarse.set_schedule(1)
        fine.upload_state(q=seed_director_field(seed, n_fine, n_fine, n_fine))
        fine.step(warmup)
        log(f"fine warmup {warmup} steps done, umax {fine.diagnostics()['umax']:.4f}")
        xs, ys, _ = generate_pairs(fine, coarse, n_fine, samples, spacing)
        # held-out pairs from a second trajectory
        fine.upload_state(q=seed_director_field(seed + 1, n_fine, n_fine, n_fine))
        fine.step(warmup)
        hx, hy, hbase = generate_pairs(fine, coarse, n_fine, heldout, spacing)
        t_data = time.perf_counter() - t0
        log(f"generated {samples} + {heldout} pairs in {t_data:.1f} s")

        X = np.stack(xs)
        Y = np.stack(ys)
        in_scale = 1.0 / np.maximum(X.std(axis=(0, 2)), 1e-12)
        out_scale = 1.0 / np.maximum(Y.std(axis=(0, 2)), 1e-12)
        with ClosureNet(nc, nc, nc, seed=seed) as net:
            net.set_normalization(in_scale, out_scale)
            for x, y in zip(xs, ys):
                net.add_sample(x, y)
            t1 = time.perf_counter()
            curve = []
            chunk = max(iters // 10, 1)
            done = 0
            while done < iters:
                k = min(chunk, iters - done)
                curve.append(net.train(k, lr, seed=done))
                done += k
                log(f"train {done}/{iters}: mean normalised loss {curve[-1]:.4f}")
            t_train = time.perf_counter() - t1

            one_no, one_cl = [], []
            for x, y, (target, qc, _) in zip(hx, hy, hbase):
                pred = net.predict(x)
                one_no.append(rel_l2(qc, target))
                one_cl.append(rel_l2(qc + pred, target))
            ms_fwd = net.time_forward(20)
            # rollout on the held-out trajectory (continues after the held-out pairs)
            start, targets = fine_targets(fine, n_fine, roll_steps, roll_every)
            err_no = rollout(coarse, None, start, targets, roll_every)
            err_cl = rollout(coarse, net, start, targets, roll_every)
            # time to solution per coarse step (= two fine steps)
            fine_ms = fine.time_steps(100) / 100 * 2
            coarse_ms = coarse.time_steps(100) / 100
            reps = 50
            coarse.synchronize()
            t2 = time.perf_counter()
            for _ in range(reps):
                coarse.step(1)
                net.apply(coarse, 1.0)
            coarse.synchronize()
            coarse_cl_ms = (time.perf_counter() - t2) * 1e3 / reps

    report.update({
        "samples": samples, "heldout": heldout, "spacing": spacing, "warmup": warmup, "iters": iters, "lr": lr,
        "data_s": t_data, "train_s": t_train, "train_ms_per_iter": 1e3 * t_train / iters, "loss_curve": curve,
        "onestep_relL2_no_closure": float(np.mean(one_no)), "onestep_relL2_closure": float(np.mean(one_cl)),
        "onestep_gain": 1.0 - float(np.mean(one_cl)) / float(np.mean(one_no)),
        "rollout_every": roll_every, "rollout_relL2_no_closure": err_no, "rollout_relL2_closure": err_cl,
        "forward_ms": ms_fwd, "forward_tflops": FLOP_PER_CELL * nc ** 3 / (ms_fwd * 1e-3) / 1e12,
        "fine_fp64_ms_per_2_steps": fine_ms, "coarse_fp32_ms_per_step": coarse_ms,
        "coarse_with_closure_wall_ms_per_step": coarse_cl_ms,
        "speedup_vs_fine_fp64": fine_ms / coarse_cl_ms,
    })
    return report


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fine", type=int, default=128)
    ap.add_argument("--samples", type=int, default=48)
    ap.add_argument("--heldout", type=int, default=8)
    ap.add_argument("--iters", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--warmup", type=int, default=1500)
    a = ap.parse_args()
    rep = run(n_fine=a.fine, samples=a.samples, heldout=a.heldout, iters=a.iters, lr=a.lr, warmup=a.warmup,
              log=lambda s: print(s, flush=True))
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
