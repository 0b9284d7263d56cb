"""Python host wrapper over the C ABI of liblbman3d.so (include/lbman3d.h).

``DeviceSolver`` mirrors the reference's device API (``DeviceSolver<BC>`` +
``DeviceFields``, src/device_solver.h:9-18, src/device_fields.h:23-80):
upload the initial state, run ``QTensorStep``/``LBMStep``/``Step``, copy the
fields back. Every call goes to the HIP library; if the library is missing
this module raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import LbmConfig, LbmDiag, LBM_OK

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbman3d.so")
_D = C.POINTER(C.c_double)
_lib = None

# Every symbol declared in include/lbman3d.h, with its ctypes signature.
SIGNATURES = {
    "lbm_default_config": (C.c_int, [C.POINTER(LbmConfig), C.c_int32, C.c_int32, C.c_int32]),
    "lbm_set_preset": (C.c_int, [C.POINTER(LbmConfig), C.c_char_p]),
    "lbm_create": (C.c_int, [C.POINTER(LbmConfig), C.POINTER(C.c_void_p)]),
    "lbm_destroy": (C.c_int, [C.c_void_p]),
    "lbm_last_error": (C.c_char_p, []),
    "lbm_device_info": (C.c_int, [C.c_char_p, C.c_size_t]),
    "lbm_get_config": (C.c_int, [C.c_void_p, C.POINTER(LbmConfig)]),
    "lbm_upload_state": (C.c_int, [C.c_void_p, _D, _D, _D, _D, _D]),
    "lbm_upload_populations": (C.c_int, [C.c_void_p, _D]),
    "lbm_upload_force": (C.c_int, [C.c_void_p, _D, _D, _D]),
    "lbm_download_state": (C.c_int, [C.c_void_p, _D, _D, _D, _D, _D]),
    "lbm_download_populations": (C.c_int, [C.c_void_p, _D]),
    "lbm_download_force": (C.c_int, [C.c_void_p, _D, _D, _D]),
    "lbm_qtensor_step": (C.c_int, [C.c_void_p]),
    "lbm_lbm_step": (C.c_int, [C.c_void_p]),
    "lbm_step": (C.c_int, [C.c_void_p, C.c_int32]),
    "lbm_synchronize": (C.c_int, [C.c_void_p]),
    "lbm_diagnostics": (C.c_int, [C.c_void_p, C.POINTER(LbmDiag)]),
    "lbm_order_director": (C.c_int, [C.c_void_p, _D, _D]),
    "lbm_order_director_noalloc": (C.c_int, [C.c_void_p, _D, _D]),
    "lbm_time_steps": (C.c_int, [C.c_void_p, C.c_int32, _D]),
    "lbm_profile_kernels": (C.c_int, [C.c_void_p, C.c_int32, _D, C.c_int32, C.c_char_p, C.c_size_t,
                                      C.POINTER(C.c_int32)]),
    "lbm_set_schedule": (C.c_int, [C.c_void_p, C.c_int32]),
    "lbm_get_schedule": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "lbm_tile_count": (C.c_int, []),
    "lbm_tile_name": (C.c_int, [C.c_int32, C.c_char_p, C.c_size_t]),
    "lbm_autotune": (C.c_int, [C.c_void_p, C.c_int32, _D, C.c_int32, C.POINTER(C.c_int32)]),
    "lbm_seed_director_field": (C.c_int, [C.c_uint64, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                          C.c_double, _D]),
    "lbm_bandwidth_probe": (C.c_int, [C.c_size_t, C.c_int32, _D, _D]),
    "lbm_debug_slot_table": (C.c_int, [C.POINTER(LbmConfig), C.POINTER(C.c_uint16)]),
    "lbm_debug_shell_region": (C.c_int, [C.POINTER(LbmConfig), C.c_int32, C.POINTER(C.c_int32), C.c_int64,
                                         C.POINTER(C.c_int64)]),
    "lbm_debug_tile_map": (C.c_int, [C.c_int32] * 5 + [C.POINTER(C.c_int32)]),
    "lbm_batch_create": (C.c_int, [C.POINTER(LbmConfig), C.c_int32, C.POINTER(C.c_void_p)]),
    "lbm_batch_destroy": (C.c_int, [C.c_void_p]),
    "lbm_batch_upload_state": (C.c_int, [C.c_void_p, C.c_int32, _D, _D, _D, _D, _D]),
    "lbm_batch_download_state": (C.c_int, [C.c_void_p, C.c_int32, _D, _D, _D, _D, _D]),
    "lbm_batch_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "lbm_batch_time_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _D]),
    "lbm_batch_diagnostics": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(LbmDiag)]),
    "lbm_batch_tile_count": (C.c_int, []),
    "lbm_batch_tile_name": (C.c_int, [C.c_int32, C.c_char_p, C.c_size_t]),
    "lbm_batch_set_tile": (C.c_int, [C.c_void_p, C.c_int32]),
    "lbm_batch_get_tile": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "lbm_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, _D, C.c_int32, C.POINTER(C.c_int32)]),
    "lbm_batch_set_grid_limit": (C.c_int, [C.c_void_p, C.c_int32]),
    "lbm_closure_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                                     C.POINTER(C.c_void_p)]),
    "lbm_closure_destroy": (C.c_int, [C.c_void_p]),
    "lbm_closure_param_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "lbm_closure_get_params": (C.c_int, [C.c_void_p, _D]),
    "lbm_closure_set_params": (C.c_int, [C.c_void_p, _D]),
    "lbm_closure_set_normalization": (C.c_int, [C.c_void_p, _D, _D]),
    "lbm_closure_add_sample": (C.c_int, [C.c_void_p, _D, _D, C.POINTER(C.c_int32)]),
    "lbm_closure_clear_samples": (C.c_int, [C.c_void_p]),
    "lbm_closure_train": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_uint64, _D]),
    "lbm_closure_loss_and_grad": (C.c_int, [C.c_void_p, _D, _D, _D, _D]),
    "lbm_closure_predict": (C.c_int, [C.c_void_p, _D, _D]),
    "lbm_closure_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double]),
    "lbm_closure_time_forward": (C.c_int, [C.c_void_p, C.c_int32, _D]),
    "lbm_schedule_name": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_size_t]),
    "lbm_closure_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_double, _D]),
    "lbm_debug_mfma_bf16": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "lbm_defect_faces": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint8), C.POINTER(C.c_uint8), C.POINTER(C.c_uint8)]),
}


class LbmError(RuntimeError):
    """A C-ABI call returned a negative status (mirrors the reference's
    std::runtime_error convention, src/cuda/cuda_utils.h:7-12)."""

    def __init__(self, code, msg):
        super().__init__(f"lbman3d error {code}: {msg}")
        self.code = code


def lib():
    """Load liblbman3d.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m lbman3d_amd.build` (or "
                               "__graft_entry__.build()) first; there is no fallback path")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != LBM_OK:
        raise LbmError(code, lib().lbm_last_error().decode())
    return code


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "arrays must be contiguous float64"
    return a.ctypes.data_as(_D)


def device_info() -> str:
    buf = C.create_string_buffer(512)
    check(lib().lbm_device_info(buf, 512))
    return buf.value.decode()


def seed_director_field(seed, nx, ny, nz, s0=0.5, noise=0.01) -> np.ndarray:
    """The library's host-side seeded random-director Q (5, V)."""
    q = np.empty(5 * nx * ny * nz)
    check(lib().lbm_seed_director_field(seed, nx, ny, nz, s0, noise, _p(q)))
    return q.reshape(5, -1)


def bandwidth_probe(nbytes=1 << 30, reps=20):
    """Achievable HBM bandwidth of a plain copy kernel: (GB/s, ms per copy)."""
    g, ms = C.c_double(), C.c_double()
    check(lib().lbm_bandwidth_probe(nbytes, reps, C.byref(g), C.byref(ms)))
    return g.value, ms.value


def slot_table(cfg: LbmConfig) -> np.ndarray:
    out = (C.c_uint16 * (64 * 15))()
    check(lib().lbm_debug_slot_table(C.byref(cfg), out))
    return np.array(out[:], dtype=np.uint16).reshape(64, 15)


def shell_region(cfg: LbmConfig, w: int) -> np.ndarray:
    """(count, 3) array of the (x, y, z) cells within w of an open face, in
    the wall-shell passes' order (host-only)."""
    n = C.c_int64()
    check(lib().lbm_debug_shell_region(C.byref(cfg), w, None, 0, C.byref(n)))
    out = (C.c_int32 * max(1, 3 * n.value))()
    check(lib().lbm_debug_shell_region(C.byref(cfg), w, out, n.value, C.byref(n)))
    return np.array(out[:3 * n.value], dtype=np.int32).reshape(-1, 3)


class DeviceSolver:
    """One simulation on the GPU (DeviceFields + DeviceSolver<BC>)."""

    def __init__(self, cfg: LbmConfig):
        self.cfg = cfg
        self.V = cfg.nx * cfg.ny * cfg.nz
        h = C.c_void_p()
        check(lib().lbm_create(C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- state ---------------------------------------------------------------
    def upload_state(self, rho=None, u=None, q=None):
        arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        rho, q = arr(rho), arr(q)
        u = [arr(u[a]) for a in range(3)] if u is not None else [None] * 3
        check(lib().lbm_upload_state(self.h, _p(rho), _p(u[0]), _p(u[1]), _p(u[2]), _p(q)))

    def upload_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        check(lib().lbm_upload_populations(self.h, _p(f)))

    def upload_force(self, F):
        F = [np.ascontiguousarray(F[a], dtype=np.float64).reshape(-1) for a in range(3)]
        check(lib().lbm_upload_force(self.h, *[_p(x) for x in F]))

    def state(self):
        V = self.V
        rho, ux, uy, uz = (np.empty(V) for _ in range(4))
        q = np.empty(5 * V)
        check(lib().lbm_download_state(self.h, _p(rho), _p(ux), _p(uy), _p(uz), _p(q)))
        return {"rho": rho, "u": np.stack([ux, uy, uz]), "q": q.reshape(5, V)}

    def force(self):
        F = [np.empty(self.V) for _ in range(3)]
        check(lib().lbm_download_force(self.h, *[_p(x) for x in F]))
        return np.stack(F)

    def populations(self):
        f = np.empty(15 * self.V)
        check(lib().lbm_download_populations(self.h, _p(f)))
        return f.reshape(15, self.V)

    # -- stepping -------------------------------------------------------------
    def qtensor_step(self):
        check(lib().lbm_qtensor_step(self.h))

    def lbm_step(self):
        check(lib().lbm_lbm_step(self.h))

    def step(self, n=1):
        check(lib().lbm_step(self.h, n))

    def synchronize(self):
        check(lib().lbm_synchronize(self.h))

    # -- diagnostics / timing ---------------------------------------------------
    def diagnostics(self) -> dict:
        d = LbmDiag()
        check(lib().lbm_diagnostics(self.h, C.byref(d)))
        return d.as_dict()

    def order_director(self, noalloc=False):
        """Order parameter and director per cell; noalloc uses the entry
        point that stages in a fixed device array (grids up to 2^19 cells)."""
        o, d = np.empty(self.V), np.empty(3 * self.V)
        fn = lib().lbm_order_director_noalloc if noalloc else lib().lbm_order_director
        check(fn(self.h, _p(o), _p(d)))
        return o, d.reshape(self.V, 3)

    def time_steps(self, n) -> float:
        ms = C.c_double()
        check(lib().lbm_time_steps(self.h, n, C.byref(ms)))
        return ms.value

    def profile_kernels(self, n) -> dict:
        ms = (C.c_double * 16)()
        names = C.create_string_buffer(1024)
        cnt = C.c_int32()
        check(lib().lbm_profile_kernels(self.h, n, ms, 16, names, 1024, C.byref(cnt)))
        keys = names.value.decode().split(";") if names.value else []
        return {k: ms[i] for i, k in enumerate(keys[:cnt.value])}

    def set_schedule(self, s: int):
        check(lib().lbm_set_schedule(self.h, s))

    def schedule_name(self, s: int) -> str:
        """Name of schedule s for this handle's precision."""
        buf = C.create_string_buffer(128)
        check(lib().lbm_schedule_name(self.h, s, buf, 128))
        return buf.value.decode()

    def get_schedule(self) -> int:
        s = C.c_int32()
        check(lib().lbm_get_schedule(self.h, C.byref(s)))
        return s.value

    def autotune(self, n=10) -> dict:
        ms = (C.c_double * 64)()
        cnt = C.c_int32()
        check(lib().lbm_autotune(self.h, n, ms, 64, C.byref(cnt)))
        return {self.schedule_name(i + 1): ms[i] for i in range(cnt.value)}


def tile_count() -> int:
    return lib().lbm_tile_count()


def batch_tile_name(t: int) -> str:
    buf = C.create_string_buffer(64)
    check(lib().lbm_batch_tile_name(t, buf, 64))
    return buf.value.decode()


def batch_tile_index(name: str) -> int:
    """Catalogue index of a batched tile shape by name."""
    for t in range(lib().lbm_batch_tile_count()):
        if batch_tile_name(t) == name:
            return t
    raise KeyError(name)


def schedule_name(s: int) -> str:
    buf = C.create_string_buffer(128)
    check(lib().lbm_tile_name(s, buf, 128))
    return buf.value.decode()


class BatchSolver:
    """Many independent FP32 simulations stepped by one launch per step
    (lbm_batch_*). Every config must be FP32_SHIFTED."""

    def __init__(self, cfgs):
        self.cfgs = list(cfgs)
        arr = (LbmConfig * len(self.cfgs))(*self.cfgs)
        h = C.c_void_p()
        check(lib().lbm_batch_create(arr, len(self.cfgs), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().lbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _V(self, i):
        c = self.cfgs[i]
        return c.nx * c.ny * c.nz

    def upload_state(self, i, rho=None, u=None, q=None):
        arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        rho, q = arr(rho), arr(q)
        u = [arr(u[a]) for a in range(3)] if u is not None else [None] * 3
        check(lib().lbm_batch_upload_state(self.h, i, _p(rho), _p(u[0]), _p(u[1]), _p(u[2]), _p(q)))

    def state(self, i):
        V = self._V(i)
        rho, ux, uy, uz = (np.empty(V) for _ in range(4))
        q = np.empty(5 * V)
        check(lib().lbm_batch_download_state(self.h, i, _p(rho), _p(ux), _p(uy), _p(uz), _p(q)))
        return {"rho": rho, "u": np.stack([ux, uy, uz]), "q": q.reshape(5, V)}

    def diagnostics(self, i) -> dict:
        d = LbmDiag()
        check(lib().lbm_batch_diagnostics(self.h, i, C.byref(d)))
        return d.as_dict()

    def step(self, n=1, use_graph=True):
        check(lib().lbm_batch_step(self.h, n, 1 if use_graph else 0))

    def time_steps(self, n, use_graph=True) -> float:
        ms = C.c_double()
        check(lib().lbm_batch_time_steps(self.h, n, 1 if use_graph else 0, C.byref(ms)))
        return ms.value

    def set_tile(self, t):
        check(lib().lbm_batch_set_tile(self.h, t))

    def get_tile(self) -> int:
        t = C.c_int32()
        check(lib().lbm_batch_get_tile(self.h, C.byref(t)))
        return t.value

    def set_grid_limit(self, max_blocks: int):
        """Cap the persistent grid (0: automatic)."""
        check(lib().lbm_batch_set_grid_limit(self.h, max_blocks))

    def autotune(self, n=5) -> list:
        ms = (C.c_double * 16)()
        cnt = C.c_int32()
        check(lib().lbm_batch_autotune(self.h, n, ms, 16, C.byref(cnt)))
        return [ms[i] for i in range(cnt.value)]
