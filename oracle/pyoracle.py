"""ctypes front end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the lbman3d_amd product package. The
oracle restates the reference CPU step (see oracle/lbm_oracle.cpp for the
file:line citations of every function).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from lbman3d_amd.abi import LbmConfig, LbmParams, LbmWall

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None
_D = C.POINTER(C.c_double)


def build(force: bool = False) -> str:
    """Compile oracle/liboracle.so with g++ (FP64, no FMA contraction)."""
    src = os.path.join(HERE, "lbm_oracle.cpp")
    hdr = os.path.join(HERE, "..", "include", "lbman3d.h")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(
            os.path.getmtime(src), os.path.getmtime(hdr)):
        tmp = LIB_PATH + ".tmp"
        subprocess.check_call(["g++", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
                               "-std=c++17", "-o", tmp, src])
        os.replace(tmp, LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(LbmConfig), C.c_int]
        for name in ("orc_destroy", "orc_init_feq", "orc_qtensor_step", "orc_lbm_step"):
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = None
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_threads.argtypes = [C.c_void_p, C.c_int]
        L.orc_set_fields.argtypes = [C.c_void_p] + [_D] * 5
        L.orc_get_fields.argtypes = [C.c_void_p] + [_D] * 5
        L.orc_set_force.argtypes = [C.c_void_p] + [_D] * 3
        L.orc_get_force.argtypes = [C.c_void_p] + [_D] * 3
        L.orc_set_populations.argtypes = [C.c_void_p, _D]
        L.orc_get_populations.argtypes = [C.c_void_p, _D]
        L.orc_get_stress.argtypes = [C.c_void_p, _D, _D]
        L.orc_free_energy.argtypes = [C.c_void_p]
        L.orc_free_energy.restype = C.c_double
        L.orc_log_sums.argtypes = [C.c_void_p, _D]
        L.orc_order_director.argtypes = [C.c_void_p, _D, _D]
        L.orc_pointwise_collide.argtypes = [C.POINTER(LbmParams), C.c_double, _D, _D, _D, _D, _D, _D]
        L.orc_moments.argtypes = [C.POINTER(LbmParams), _D, _D, _D]
        L.orc_pointwise_q.argtypes = [C.POINTER(LbmParams), C.c_int, _D, _D, _D, _D, _D, _D, _D, _D]
        L.orc_q_derivs.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _D]
        L.orc_velocity_gradient.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _D]
        L.orc_stress_divergence.argtypes = [C.c_void_p, _D, _D, C.c_int, C.c_int, C.c_int, _D]
        L.orc_anchored_q.argtypes = [C.POINTER(LbmWall), _D]
        L.orc_lattice.argtypes = [C.POINTER(C.c_int)] * 3 + [_D] + [C.POINTER(C.c_int)] * 2
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_D)


def _arr(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One FP64 CPU simulation, reference semantics (push streaming, AoS)."""

    def __init__(self, cfg: LbmConfig, threads: int = 1):
        self.cfg = cfg
        self.V = cfg.nx * cfg.ny * cfg.nz
        self.h = lib().orc_create(C.byref(cfg), threads)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_threads(self, n):
        lib().orc_set_threads(self.h, n)

    def set_state(self, rho=None, u=None, q=None):
        """Set fields and initialise f = Feq(rho, u) (LbmSolver::Initialize)."""
        rho, q = _arr(rho), _arr(q)
        u = None if u is None else [_arr(u[a]) for a in range(3)]
        lib().orc_set_fields(self.h, _p(rho), *( [_p(u[a]) for a in range(3)] if u else [None] * 3),
                             _p(q))
        lib().orc_init_feq(self.h)

    def set_q(self, q):
        lib().orc_set_fields(self.h, None, None, None, None, _p(_arr(q)))

    def set_force(self, F):
        F = [_arr(F[a]) for a in range(3)]
        lib().orc_set_force(self.h, *[_p(x) for x in F])

    def set_populations(self, f):
        lib().orc_set_populations(self.h, _p(_arr(f)))

    def qtensor_step(self):
        lib().orc_qtensor_step(self.h)

    def lbm_step(self):
        lib().orc_lbm_step(self.h)

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def state(self):
        V = self.V
        rho, ux, uy, uz = (np.empty(V) for _ in range(4))
        q = np.empty(5 * V)
        lib().orc_get_fields(self.h, _p(rho), _p(ux), _p(uy), _p(uz), _p(q))
        return {"rho": rho, "u": np.stack([ux, uy, uz]), "q": q.reshape(5, V)}

    def force(self):
        F = [np.empty(self.V) for _ in range(3)]
        lib().orc_get_force(self.h, *[_p(x) for x in F])
        return np.stack(F)

    def populations(self):
        f = np.empty(15 * self.V)
        lib().orc_get_populations(self.h, _p(f))
        return f.reshape(15, self.V)

    def stress(self):
        s, t = np.empty(5 * self.V), np.empty(3 * self.V)
        lib().orc_get_stress(self.h, _p(s), _p(t))
        return s.reshape(5, self.V), t.reshape(3, self.V)

    def free_energy(self):
        return lib().orc_free_energy(self.h)

    def log_sums(self):
        out = np.empty(8)
        lib().orc_log_sums(self.h, _p(out))
        return dict(zip(("mass", "px", "py", "pz", "kinetic", "du2", "u2", "umax"), out))

    def order_director(self):
        o, d = np.empty(self.V), np.empty(3 * self.V)
        lib().orc_order_director(self.h, _p(o), _p(d))
        return o, d.reshape(self.V, 3)

    def q_derivs(self, x, y, z):
        out = np.empty(35)
        lib().orc_q_derivs(self.h, x, y, z, _p(out))
        return out.reshape(5, 7)

    def velocity_gradient(self, x, y, z):
        out = np.empty(9)
        lib().orc_velocity_gradient(self.h, x, y, z, _p(out))
        return out.reshape(3, 3)

    def stress_divergence(self, sigma, tau, x, y, z):
        out = np.empty(3)
        lib().orc_stress_divergence(self.h, _p(_arr(sigma)), _p(_arr(tau)), x, y, z, _p(out))
        return out


def pointwise_collide(params: LbmParams, rho, u, F, f):
    feq, forcing, fstar = np.empty(15), np.empty(15), np.empty(15)
    lib().orc_pointwise_collide(C.byref(params), rho, _p(_arr(u)), _p(_arr(F)), _p(_arr(f)),
                                _p(feq), _p(forcing), _p(fstar))
    return feq, forcing, fstar


def moments(params: LbmParams, f, F):
    out = np.empty(4)
    lib().orc_moments(C.byref(params), _p(_arr(f)), _p(_arr(F)), _p(out))
    return out


def pointwise_q(params: LbmParams, scheme, Q, u, derivs, grad):
    qn, sg, ta, fe = np.empty(5), np.empty(5), np.empty(3), np.empty(3)
    lib().orc_pointwise_q(C.byref(params), scheme, _p(_arr(Q)), _p(_arr(u)),
                          _p(_arr(np.asarray(derivs).reshape(35))), _p(_arr(np.asarray(grad).reshape(9))),
                          _p(qn), _p(sg), _p(ta), _p(fe))
    return qn, sg, ta, fe


def anchored_q(wall: LbmWall):
    out = np.empty(5)
    lib().orc_anchored_q(C.byref(wall), _p(out))
    return out


def lattice():
    ex, ey, ez, opp = ((C.c_int * 15)() for _ in range(4))
    spec = (C.c_int * 45)()
    w = np.empty(15)
    lib().orc_lattice(ex, ey, ez, _p(w), opp, spec)
    return (np.array(ex[:]), np.array(ey[:]), np.array(ez[:]), w, np.array(opp[:]),
            np.array(spec[:]).reshape(3, 15))


# ---- the reference's own solver core (oracle/_ref, built by build_ref.sh) ----

def ref_binary(scenario: str) -> str:
    return os.path.join(REF_DIR, f"ref_{scenario}")


def run_ref(scenario: str, bc: str, mode: str, force_model: int, steps: int, dims, q,
            fields=None, force=None, workdir=None):
    """Run oracle/_ref/ref_<scenario> and return its outputs as a dict."""
    import tempfile
    nx, ny, nz = dims
    V = nx * ny * nz
    wd = workdir or tempfile.mkdtemp(prefix="lbmref_")
    fin, fout = os.path.join(wd, "in.bin"), os.path.join(wd, "out.bin")
    with open(fin, "wb") as fp:
        np.array([nx, ny, nz, 1 if fields is not None else 0, 1 if force is not None else 0],
                 dtype=np.int32).tofile(fp)
        np.ascontiguousarray(q, dtype=np.float64).reshape(-1).tofile(fp)
        if fields is not None:
            np.ascontiguousarray(fields, dtype=np.float64).reshape(-1).tofile(fp)
        if force is not None:
            np.ascontiguousarray(force, dtype=np.float64).reshape(-1).tofile(fp)
    r = subprocess.run([ref_binary(scenario), bc, mode, str(force_model), str(steps), fin, fout],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ref_{scenario} failed: {r.stderr}")
    raw = np.fromfile(fout, dtype=np.float64)
    o = 0

    def take(n):
        nonlocal o
        v = raw[o:o + n]
        o += n
        return v
    out = {"rho": take(V), "u": take(3 * V).reshape(3, V), "q": take(5 * V).reshape(5, V),
           "F": take(3 * V).reshape(3, V), "f": take(15 * V).reshape(15, V),
           "free_energy": take(1)[0], "log": take(8), "stderr": r.stderr}
    return out


def run_ref_defects(bc: str, dims, q, workdir=None):
    """Run oracle/_ref/ref_defects (the reference's QtensorToOrderDirector and
    DefectFinder<BC>, built by build_ref.sh for a 14 x 12 x 10 grid) on Q
    (5 x V) and return the order, director [V][3], the three face arrays and
    the disclination paths (each an n x 3 array of points)."""
    import tempfile
    nx, ny, nz = dims
    V = nx * ny * nz
    wd = workdir or tempfile.mkdtemp(prefix="lbmdef_")
    fin, fout = os.path.join(wd, "in.bin"), os.path.join(wd, "out.bin")
    with open(fin, "wb") as fp:
        np.array([nx, ny, nz], dtype=np.int32).tofile(fp)
        np.ascontiguousarray(q, dtype=np.float64).reshape(-1).tofile(fp)
    r = subprocess.run([ref_binary("defects"), bc, fin, fout], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"ref_defects failed: {r.stderr}")
    raw = open(fout, "rb").read()
    o = 0

    def take(dtype, n):
        nonlocal o
        a = np.frombuffer(raw, dtype=dtype, count=n, offset=o)
        o += a.nbytes
        return a.copy()
    order = take(np.float64, V)
    director = take(np.float64, 3 * V).reshape(V, 3)
    faces = tuple(take(np.uint8, n) for n in ((nx * (ny - 1) * (nz - 1)), (nx - 1) * ny * (nz - 1),
                                              (nx - 1) * (ny - 1) * nz))
    paths = []
    for _ in range(int(take(np.int64, 1)[0])):
        n = int(take(np.int64, 1)[0])
        paths.append(take(np.float64, 3 * n).reshape(n, 3))
    assert o == len(raw)
    return {"order": order, "director": director, "faces": faces, "paths": paths}
