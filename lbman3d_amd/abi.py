"""ctypes mirror of include/lbman3d.h plus host-side configuration helpers.

The names follow the reference's compile-time API surface:
``Periodic``/``Neumann``/``Anchoring``/``NoSlip``/``MovingWall``/
``SpecularReflection`` tags and ``WallSpec`` (src/boundary.h:31-76), the
presets ``FullyPeriodicConfig``/``ChannelConfig``/``SlitConfig``/
``SlitNoSlipConfig`` (src/boundary.h:105-124, src/sim_config.h:19-45),
``Params`` defaults (src/params.h:14-112) and ``Advection``
(src/qtensor_types.h:22).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
import math

import numpy as np

LBM_OK = 0
LBM_ERR_INVALID = -1
LBM_ERR_UNSUPPORTED = -2
LBM_ERR_HIP = -3
LBM_ERR_STATE = -4

Q_PERIODIC, Q_NEUMANN, Q_ANCHORING = 0, 1, 2
U_PERIODIC, U_NOSLIP, U_MOVING, U_SPECULAR = 0, 1, 2, 3
XLO, XHI, YLO, YHI, ZLO, ZHI = range(6)
ADV_CENTRED, ADV_UPWIND = 0, 1
FP64, FP32_SHIFTED = 0, 1
FORCE_FULL, FORCE_ERICKSEN_ONLY = 0, 1


class LbmWall(C.Structure):
    _fields_ = [("qbc", C.c_int32), ("ubc", C.c_int32), ("anchor_s", C.c_double),
                ("anchor_theta", C.c_double), ("anchor_phi", C.c_double),
                ("wall_u", C.c_double * 3)]


class LbmParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("dt", "density", "tauf", "L", "A", "B", "C", "lam", "gamma", "alpha", "mu")]


class LbmConfig(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("advection", C.c_int32), ("precision", C.c_int32), ("force_model", C.c_int32),
                ("walls", LbmWall * 6), ("p", LbmParams)]


class LbmDiag(C.Structure):
    _fields_ = [("mass", C.c_double), ("px", C.c_double), ("py", C.c_double), ("pz", C.c_double),
                ("kinetic", C.c_double), ("du2", C.c_double), ("u2", C.c_double),
                ("umax", C.c_double), ("free_energy", C.c_double), ("nonfinite", C.c_int32),
                ("_pad", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "_pad"}


# ---- boundary tags -----------------------------------------------------------

@dataclass(frozen=True)
class Periodic:
    pass


@dataclass(frozen=True)
class Neumann:
    pass


@dataclass(frozen=True)
class NoSlip:
    pass


@dataclass(frozen=True)
class SpecularReflection:
    pass


@dataclass(frozen=True)
class Anchoring:
    """Strong anchoring at order S and director angles (theta, phi) in radians."""
    s: float = 1.0
    theta: float = 0.0
    phi: float = 0.0


@dataclass(frozen=True)
class MovingWall:
    Ux: float = 0.0
    Uy: float = 0.0
    Uz: float = 0.0


@dataclass(frozen=True)
class WallSpec:
    QBC: object
    UBC: object


def _wall_struct(ws: WallSpec) -> LbmWall:
    w = LbmWall()
    q, u = ws.QBC, ws.UBC
    if isinstance(q, Periodic):
        w.qbc = Q_PERIODIC
    elif isinstance(q, Neumann):
        w.qbc = Q_NEUMANN
    elif isinstance(q, Anchoring):
        w.qbc = Q_ANCHORING
        w.anchor_s, w.anchor_theta, w.anchor_phi = q.s, q.theta, q.phi
    else:
        raise TypeError(f"not a Q boundary tag: {q!r}")
    if isinstance(u, Periodic):
        w.ubc = U_PERIODIC
    elif isinstance(u, NoSlip):
        w.ubc = U_NOSLIP
    elif isinstance(u, SpecularReflection):
        w.ubc = U_SPECULAR
    elif isinstance(u, MovingWall):
        w.ubc = U_MOVING
        w.wall_u[0], w.wall_u[1], w.wall_u[2] = u.Ux, u.Uy, u.Uz
    else:
        raise TypeError(f"not a velocity boundary tag: {u!r}")
    return w


@dataclass(frozen=True)
class BCConfig:
    """Six faces, XLo..ZHi, each a WallSpec (src/boundary.h:82-98)."""
    XLo: WallSpec = WallSpec(Neumann(), NoSlip())
    XHi: WallSpec = WallSpec(Neumann(), NoSlip())
    YLo: WallSpec = WallSpec(Neumann(), NoSlip())
    YHi: WallSpec = WallSpec(Neumann(), NoSlip())
    ZLo: WallSpec = WallSpec(Neumann(), NoSlip())
    ZHi: WallSpec = WallSpec(Neumann(), NoSlip())
    name: str = "Custom"

    def faces(self):
        return (self.XLo, self.XHi, self.YLo, self.YHi, self.ZLo, self.ZHi)


_PP = WallSpec(Periodic(), Periodic())
FullyPeriodicConfig = BCConfig(_PP, _PP, _PP, _PP, _PP, _PP, "FullyPeriodic")
PeriodicBC = FullyPeriodicConfig
ChannelConfig = BCConfig(WallSpec(Neumann(), NoSlip()), WallSpec(Neumann(), NoSlip()),
                         WallSpec(Neumann(), NoSlip()), WallSpec(Neumann(), NoSlip()),
                         _PP, _PP, "Channel")
SlitConfig = BCConfig(WallSpec(Neumann(), SpecularReflection()),
                      WallSpec(Neumann(), SpecularReflection()), _PP, _PP, _PP, _PP, "SlitFreeSlip")
SlitNoSlipConfig = BCConfig(WallSpec(Neumann(), NoSlip()), WallSpec(Neumann(), NoSlip()),
                            _PP, _PP, _PP, _PP, "SlitNoSlip")
PRESETS = {c.name: c for c in (FullyPeriodicConfig, ChannelConfig, SlitConfig, SlitNoSlipConfig)}


# ---- physical constants (namespace Params, src/params.h) --------------------

@dataclass
class Params:
    DT: float = 1.0
    kDensity: float = 1.0
    TAUF: float = 2.5
    L: float = 0.06
    A: float = 0.0
    B: float = -0.3
    C: float = 0.3
    LAMBDA: float = 0.3
    GAMMA: float = 0.34
    ALPHA: float = 0.023
    MU: float = 0.0

    def struct(self) -> LbmParams:
        p = LbmParams()
        p.dt, p.density, p.tauf, p.L = self.DT, self.kDensity, self.TAUF, self.L
        p.A, p.B, p.C = self.A, self.B, self.C
        p.lam, p.gamma, p.alpha, p.mu = self.LAMBDA, self.GAMMA, self.ALPHA, self.MU
        return p


def make_config(nx, ny, nz, bc: BCConfig = FullyPeriodicConfig, params: Params | None = None,
                advection=ADV_CENTRED, precision=FP64, force_model=FORCE_FULL) -> LbmConfig:
    cfg = LbmConfig()
    cfg.nx, cfg.ny, cfg.nz = nx, ny, nz
    cfg.advection, cfg.precision, cfg.force_model = advection, precision, force_model
    for i, ws in enumerate(bc.faces()):
        cfg.walls[i] = _wall_struct(ws)
    cfg.p = (params or Params()).struct()
    return cfg


def config_copy(cfg: LbmConfig, **changes) -> LbmConfig:
    out = LbmConfig()
    C.memmove(C.byref(out), C.byref(cfg), C.sizeof(LbmConfig))
    for k, v in changes.items():
        setattr(out, k, v)
    return out


# ---- seeded initial state -------------------------------------------------

_M64 = (1 << 64) - 1


def _splitmix64_np(x: np.ndarray) -> np.ndarray:
    z = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(_M64)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(_M64)
    return z ^ (z >> np.uint64(31))


def uniform_stream(seed: int, V: int, k: int) -> np.ndarray:
    """U[0,1) per cell from splitmix64 of the key (seed, cell, k): the same
    counter-based stream as lbm_seed_director_field."""
    with np.errstate(over="ignore"):
        cells = np.arange(V, dtype=np.uint64)
        key = (np.uint64(seed & _M64) * np.uint64(0x100000001B3)) & np.uint64(_M64)
        key = key ^ (cells * np.uint64(8) + np.uint64(k))
        r = _splitmix64_np(_splitmix64_np(key))
    return (r >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def seed_director_field(seed, nx, ny, nz, s0=0.5, noise=0.01) -> np.ndarray:
    """Numpy restatement of lbm_seed_director_field (used to cross-check the
    library's host generator): n uniform on S^2 via z = 2U-1, phi = 2 pi U,
    S = s0 + noise*(2U-1), Q = S(nn - I/3) -> (xx, xy, xz, yy, yz)."""
    V = nx * ny * nz
    u0, u1, u2 = (uniform_stream(seed, V, k) for k in range(3))
    zc = 2.0 * u0 - 1.0
    ph = 2.0 * math.pi * u1
    rr = np.sqrt(np.maximum(0.0, 1.0 - zc * zc))
    n = (rr * np.cos(ph), rr * np.sin(ph), zc)
    S = s0 + noise * (2.0 * u2 - 1.0)
    q = np.empty((5, V))
    q[0] = S * (n[0] * n[0] - 1.0 / 3.0)
    q[1] = S * (n[0] * n[1])
    q[2] = S * (n[0] * n[2])
    q[3] = S * (n[1] * n[1] - 1.0 / 3.0)
    q[4] = S * (n[1] * n[2])
    return q
