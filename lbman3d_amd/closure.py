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
fine state. The fine trajectories start from seeded random directors with
activity zeta drawn uniformly from [0.5, 2] x the params.h default (0.023);
the held-out trajectories use their own seeds and their own zeta draws. The
coarse solver of a trajectory runs at that trajectory's zeta; the network
sees only (Q, u).

The report gives, on the held-out trajectories, the one-step error with and
without the closure, a 40-coarse-step rollout with and without it, and the
time-to-solution comparison at equal error: the FP64 plain step run at
intermediate grids N^3 (N = 64, 80, 96, 112, 128, acoustic scaling for the
ratio 128/N, started from the Fourier-resampled fine state) is judged by the
same measure (resampled to 128^3 and box-filtered to 64^3) against the same
fine trajectory, and its time for the 40 coarse steps (= 80 fine steps) is
set beside the closure run's.

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

    def run(self, sim: DeviceSolver, n, alpha=1.0):
        """n steps of sim, each followed by the correction; device ms."""
        ms = C.c_double()
        check(lib().lbm_closure_run(self.h, sim.h, n, alpha, C.byref(ms)))
        return ms.value

    def time_forward(self, n=20):
        ms = C.c_double()
        check(lib().lbm_closure_time_forward(self.h, n, C.byref(ms)))
        return ms.value


# ---- fine / coarse models -------------------------------------------------

ZETA_DEFAULT = 0.023   # params.h ALPHA


def coarse_params(p: abi.Params, r: float = 2.0) -> abi.Params:
    """Constants of a lattice coarsened by the factor r under acoustic
    scaling (dx and dt both multiplied by r)."""
    return abi.Params(DT=p.DT, kDensity=p.kDensity, TAUF=0.5 + (p.TAUF - 0.5) / r, L=p.L / (r * r), A=p.A,
                      B=p.B, C=p.C, LAMBDA=p.LAMBDA, GAMMA=r * p.GAMMA, ALPHA=p.ALPHA, MU=p.MU)


def fine_config(n=128, params=None):
    return abi.make_config(n, n, n, abi.FullyPeriodicConfig, params or abi.Params(), advection=abi.ADV_UPWIND,
                           precision=abi.FP64)


def coarse_config(n=64, params=None, r=2.0, precision=abi.FP32_SHIFTED):
    return abi.make_config(n, n, n, abi.FullyPeriodicConfig, coarse_params(params or abi.Params(), r),
                           advection=abi.ADV_UPWIND, precision=precision)


def trajectories(count, first_seed, draw_seed):
    """(director seed, zeta) of `count` fine trajectories: zeta uniform in
    [0.5, 2] x the default."""
    rng = np.random.default_rng(draw_seed)
    return [(first_seed + k, float(rng.uniform(0.5, 2.0)) * ZETA_DEFAULT) for k in range(count)]


def filter2(a, n):
    """2x2x2 box filter of channel-major fields (..., n^3) -> (..., (n/2)^3)."""
    lead = a.shape[:-1]
    g = a.reshape(*lead, n // 2, 2, n // 2, 2, n // 2, 2)
    k = len(lead)
    return g.mean(axis=(k + 1, k + 3, k + 5)).reshape(*lead, (n // 2) ** 3)


def fourier_resample(a, n_in, n_out):
    """Periodic band-limited resampling of channel-major fields (..., n_in^3)
    -> (..., n_out^3): the spectrum is truncated or zero-padded to the modes
    |k| < min(n_in, n_out) / 2 on each axis (the Nyquist plane is dropped, so
    the result stays real)."""
    if n_in == n_out:
        return a.copy()
    lead = a.shape[:-1]
    g = a.reshape(-1, n_in, n_in, n_in)
    m = min(n_in, n_out) // 2
    idx_in = np.r_[0:m, n_in - m + 1:n_in]
    idx_out = np.r_[0:m, n_out - m + 1:n_out]
    out = np.empty((g.shape[0], n_out, n_out, n_out))
    for c in range(g.shape[0]):
        spec = np.fft.fftn(g[c])
        dst = np.zeros((n_out, n_out, n_out), dtype=complex)
        dst[np.ix_(idx_out, idx_out, idx_out)] = spec[np.ix_(idx_in, idx_in, idx_in)]
        out[c] = np.fft.ifftn(dst).real * (n_out / n_in) ** 3
    return out.reshape(*lead, n_out ** 3)


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
    """Fine state at the start (full and filtered) and the filtered fine Q
    every `every` coarse steps."""
    full = fine.state()
    st = {"rho": filter2(full["rho"], n), "u": filter2(full["u"], n), "q": filter2(full["q"], n)}
    targets = []
    for _ in range(steps // every):
        fine.step(2 * every)
        targets.append(filter2(fine.state()["q"], n))
    return full, st, targets


def rollout(coarse: DeviceSolver, net: ClosureNet | None, start, targets, every):
    """Coarse run from the filtered fine state (closure added after every
    step when `net` is given); relative L2 of Q against each target."""
    coarse.upload_state(rho=start["rho"], u=start["u"], q=start["q"])
    errs = []
    for t in targets:
        if net is not None:
            net.run(coarse, every, 1.0)
        else:
            coarse.step(every)
        errs.append(rel_l2(coarse.state()["q"], t))
    return errs


def plain_rollout(n, n_fine, params, full_start, target, coarse_steps):
    """FP64 plain run on an n^3 grid (acoustic scaling for n_fine / n) over
    the span of `coarse_steps` steps of the 2x-coarsened model (= 2 x that
    many fine steps), from the Fourier-resampled fine state; its Q is
    resampled to n_fine^3 and box-filtered to the closure's grid to be
    judged against the same target. Returns (relative L2, ms for the span)."""
    steps = 2 * coarse_steps * n // n_fine
    assert steps * n_fine == 2 * coarse_steps * n, "span must be a whole number of steps on this grid"
    cfg = coarse_config(n, params, r=n_fine / n, precision=abi.FP64)
    with DeviceSolver(cfg) as s:
        s.upload_state(rho=fourier_resample(full_start["rho"], n_fine, n),
                       u=fourier_resample(full_start["u"], n_fine, n),
                       q=fourier_resample(full_start["q"], n_fine, n))
        s.step(steps)
        q = s.state()["q"]
        err = rel_l2(filter2(fourier_resample(q, n, n_fine), n_fine), target)
        ms = s.time_steps(steps)
    return err, ms


def equal_error_time(points, err):
    """Time of the FP64 plain path at a given error, interpolated in
    (log error, log time) between the measured grids; None if no grid
    reaches it."""
    pts = sorted(points, key=lambda p: p["n"])
    best = None
    for a, b in zip(pts, pts[1:]):
        if a["err"] >= err >= b["err"] and a["err"] > b["err"] > 0:
            t = (np.log(err) - np.log(a["err"])) / (np.log(b["err"]) - np.log(a["err"]))
            best = float(np.exp(np.log(a["ms"]) + t * (np.log(b["ms"]) - np.log(a["ms"]))))
            break
    if best is None and pts and pts[0]["err"] <= err:
        best = pts[0]["ms"]
    return best


def run(n_fine=128, samples=48, train_traj=4, heldout=8, heldout_traj=2, spacing=20, warmup=1500, iters=1500,
        lr=1e-3, roll_steps=40, roll_every=5, seed=11, plain_grids=(64, 80, 96, 112, 128), log=print):
    nc = n_fine // 2
    t0 = time.perf_counter()
    p0 = abi.Params()
    train = trajectories(train_traj, 1000 + seed, seed)
    held = trajectories(heldout_traj, 2000 + seed, seed + 1)
    report = {"workload": f"closure {nc}^3 coarse vs {n_fine}^3 fine", "fine": f"{n_fine}^3 FP64 upwind",
              "coarse": f"{nc}^3 FP32 upwind", "coarse_params_at_default_zeta": coarse_params(p0).__dict__,
              "train_trajectories": [{"seed": sd, "zeta": z} for sd, z in train],
              "heldout_trajectories": [{"seed": sd, "zeta": z} for sd, z in held]}

    def model(zeta):
        p = abi.Params(ALPHA=zeta)
        fine = DeviceSolver(fine_config(n_fine, p))
        coarse = DeviceSolver(coarse_config(nc, p))
        return p, fine, coarse

    xs, ys = [], []
    per = -(-samples // train_traj)
    for sd, zeta in train:
        p, fine, coarse = model(zeta)
        with fine, coarse:
            fine.upload_state(q=seed_director_field(sd, n_fine, n_fine, n_fine))
            fine.step(warmup)
            x, y, _ = generate_pairs(fine, coarse, n_fine, min(per, samples - len(xs)), spacing)
            xs += x
            ys += y
        log(f"train trajectory seed {sd} zeta {zeta:.4f}: {len(xs)} pairs")
    t_data = time.perf_counter() - t0

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
        ms_fwd = net.time_forward(20)

        one_no, one_cl, roll_no, roll_cl, plain = [], [], [], [], []
        timing = {}
        for sd, zeta in held:
            p, fine, coarse = model(zeta)
            with fine, coarse:
                fine.upload_state(q=seed_director_field(sd, n_fine, n_fine, n_fine))
                fine.step(warmup)
                hx, hy, hbase = generate_pairs(fine, coarse, n_fine, max(heldout // heldout_traj, 1), spacing)
                for x, y, (target, qc, _) in zip(hx, hy, hbase):
                    pred = net.predict(x)
                    one_no.append(rel_l2(qc, target))
                    one_cl.append(rel_l2(qc + pred, target))
                # rollout on the held-out trajectory (continues after its pairs)
                full, start, targets = fine_targets(fine, n_fine, roll_steps, roll_every)
                roll_no.append(rollout(coarse, None, start, targets, roll_every))
                roll_cl.append(rollout(coarse, net, start, targets, roll_every))
                pts = []
                for n in plain_grids:
                    err, ms = plain_rollout(n, n_fine, p, full, targets[-1], roll_steps)
                    pts.append({"n": n, "err": err, "ms": ms})
                plain.append(pts)
                log(f"held-out seed {sd} zeta {zeta:.4f}: rollout {roll_no[-1][-1]:.3e} -> {roll_cl[-1][-1]:.3e}; "
                    f"FP64 plain {[(q['n'], round(q['err'], 5)) for q in pts]}")
                if not timing:
                    # time to solution of the 40-step span (device time, HIP events)
                    timing["fine_fp64_ms"] = fine.time_steps(2 * roll_steps)
                    timing["coarse_fp32_ms"] = coarse.time_steps(roll_steps)
                    timing["coarse_with_closure_ms"] = net.run(coarse, roll_steps, 1.0)
                    coarse.synchronize()
                    t2 = time.perf_counter()
                    for _ in range(roll_steps):
                        coarse.step(1)
                        net.apply(coarse, 1.0)
                    coarse.synchronize()
                    timing["coarse_with_closure_wall_ms_python_loop"] = (time.perf_counter() - t2) * 1e3

    roll_no_m = np.mean(np.array(roll_no), axis=0).tolist()
    roll_cl_m = np.mean(np.array(roll_cl), axis=0).tolist()
    plain_m = [{"n": n, "err": float(np.mean([pts[k]["err"] for pts in plain])),
                "ms": float(np.mean([pts[k]["ms"] for pts in plain]))} for k, n in enumerate(plain_grids)]
    closure_err = roll_cl_m[-1]
    t_eq = equal_error_time(plain_m, closure_err)
    report.update({
        "samples": len(xs), "heldout": len(one_no), "spacing": spacing, "warmup": warmup, "iters": iters, "lr": lr,
        "data_s": t_data, "train_s": t_train, "train_ms_per_iter": 1e3 * t_train / iters, "loss_curve": curve,
        "onestep_relL2_no_closure": float(np.mean(one_no)), "onestep_relL2_closure": float(np.mean(one_cl)),
        "onestep_gain": 1.0 - float(np.mean(one_cl)) / float(np.mean(one_no)),
        "rollout_every": roll_every, "rollout_relL2_no_closure": roll_no_m, "rollout_relL2_closure": roll_cl_m,
        "rollout_per_trajectory_no_closure": roll_no, "rollout_per_trajectory_closure": roll_cl,
        "forward_ms": ms_fwd, "forward_tflops": FLOP_PER_CELL * nc ** 3 / (ms_fwd * 1e-3) / 1e12,
        "span_coarse_steps": roll_steps, "timing_ms_for_span": timing,
        "fp64_plain_by_grid": plain_m,
        "fp64_plain_ms_at_closure_error": t_eq,
        "speedup_vs_fp64_plain_at_equal_error": (t_eq / timing["coarse_with_closure_ms"]) if t_eq else None,
        "speedup_vs_fine_fp64": timing["fine_fp64_ms"] / timing["coarse_with_closure_ms"],
    })
    return report


def time_training(n=64, steps=200, samples=4, seed=0):
    """Device time of the training step alone: a 64^3 network with random
    (Q, u) samples and targets, `steps` Adam steps timed after 20 untimed
    ones. Returns ms per step and the forward time."""
    rng = np.random.default_rng(seed)
    with ClosureNet(n, n, n, hidden=32, seed=1) as net:
        for _ in range(samples):
            net.add_sample(rng.normal(size=(8, n ** 3)), rng.normal(size=(5, n ** 3)) * 0.2)
        net.train(20, 1e-3)
        t0 = time.perf_counter()
        net.train(steps, 1e-3)
        dt = time.perf_counter() - t0
        return {"grid": n, "train_ms_per_step": 1e3 * dt / steps, "forward_ms": net.time_forward(50)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--time-training", type=int, default=0, metavar="STEPS",
                    help="only time STEPS training steps on random 64^3 samples")
    ap.add_argument("--fine", type=int, default=128)
    ap.add_argument("--samples", type=int, default=48)
    ap.add_argument("--heldout", type=int, default=8)
    ap.add_argument("--iters", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--warmup", type=int, default=1500)
    a = ap.parse_args()
    if a.time_training:
        print(json.dumps(time_training(steps=a.time_training)), flush=True)
        return
    rep = run(n_fine=a.fine, samples=a.samples, heldout=a.heldout, iters=a.iters, lr=a.lr, warmup=a.warmup,
              log=lambda s: print(s, flush=True))
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
