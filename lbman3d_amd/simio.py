"""The run log (lbm.log) of a simulation driven through this library.

Restates the reference's log writer (SimIO, src/sim_io.cc:19-165): the
parameter summary written once, one line per Log call with the FP64 domain
reductions, the divergence verdict, and the closing line. The reductions
come from lbm_diagnostics, computed on the device in FP64 (mass, momentum,
kinetic energy, the relative velocity change since the previous Log call,
max |u| and the nematic free energy).

Numbers are written the way the reference's "{}" format specifier writes
them when it is built with std::format (src/format_compat.h, the C++23
branch): for a double that is std::to_chars without a format, the shortest
round-trip digits in fixed or scientific notation, whichever is shorter,
fixed on a tie (so 200000.0 is "2e+05"). ``fmt_double`` restates that rule.
The {fmt} fallback of that header switches to scientific notation by the
exponent instead (200000.0 is "200000"); this writer does not follow it.
"""
from __future__ import annotations

import math
from decimal import Decimal

from . import abi

NOISE = 0.05   # src/params.h NOISE: logged, not used by the seeded generator
CS2_INV = 3.0  # src/params.h kCs2Inv


def fmt_double(x: float) -> str:
    """A double as std::to_chars(first, last, x) writes it."""
    x = float(x)
    if math.isnan(x):
        return "-nan" if math.copysign(1.0, x) < 0 else "nan"
    if math.isinf(x):
        return "-inf" if x < 0 else "inf"
    sign = "-" if math.copysign(1.0, x) < 0 else ""
    if x == 0.0:
        return sign + "0"
    # shortest round-trip digits (Python's repr) as an exact decimal
    t = Decimal(repr(abs(x))).normalize().as_tuple()
    digits = "".join(map(str, t.digits))
    n = len(digits)
    point = n + t.exponent            # decimal point position after `point` digits
    sexp = point - 1
    sci = digits[0] + ("." + digits[1:] if n > 1 else "") + ("e-" if sexp < 0 else "e+") + f"{abs(sexp):02d}"
    if point <= 0:
        fixed = "0." + "0" * (-point) + digits
    elif point >= n:
        fixed = str(int(abs(x)))      # an integral value is written exactly, not digit-padded
    else:
        fixed = digits[:point] + "." + digits[point:]
    return sign + (fixed if len(fixed) <= len(sci) else sci)


def _fmt(v) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return fmt_double(v)
    return str(v)


class SimLog:
    """lbm.log writer (SimIO). ``log`` returns False once the run diverged
    (NaN mass or momentum), as SimIO::Log does."""

    def __init__(self, path="lbm.log"):
        self.f = open(path, "w")

    def _line(self, s=""):
        self.f.write(s + "\n")

    def setup_summary(self, cfg: abi.LbmConfig, bc_name: str, backend_info: str):
        """SimIO::LogSetupSummary (src/sim_io.cc:32-75)."""
        p = cfg.p
        visc = p.density / CS2_INV * (p.tauf - 0.5 * p.dt)
        L = self._line
        L("Hybrid Lattice Boltzmann simulation for 3D active nematics\n")
        L("#" * 58)
        L("#####################   Parameters   #####################")
        L("#" * 58)
        L()
        L("--- Compute backend ---")
        L(f"  {backend_info}")
        L()
        L("--- Grid ---")
        L(f"  nx = {cfg.nx}, ny = {cfg.ny}, nz = {cfg.nz}")
        L(f"  DX = {_fmt(1.0)}, DY = {_fmt(1.0)}, DZ = {_fmt(1.0)}")
        L(f"  BC = {bc_name}")
        L()
        L("--- Time ---")
        L(f"  DT (numerical)      = {_fmt(p.dt)}")
        L()
        L("--- LBM ---")
        L(f"  TAUF                = {_fmt(p.tauf)}")
        L(f"  Kinematic Viscosity = {_fmt(visc)}")
        L(f"  omega               = {_fmt(1.0 - p.dt / p.tauf)}")
        L(f"  omega_prime         = {_fmt(p.dt / p.tauf)}")
        L(f"  omega_forcing       = {_fmt(1.0 - p.dt / 2.0 / p.tauf)}")
        L()
        L("--- Free energy ---")
        L(f"  L (Frank elasticity) = {_fmt(p.L)}")
        L(f"  A                    = {_fmt(p.A)}")
        L(f"  B                    = {_fmt(p.B)}")
        L(f"  C                    = {_fmt(p.C)}")
        L()
        L("--- Q-tensor dynamics ---")
        L(f"  GAMMA (rot. viscosity^-1) = {_fmt(p.gamma)}")
        L(f"  LAMBDA (flow-aligning)    = {_fmt(p.lam)}")
        L(f"  NOISE                     = {_fmt(NOISE)}")
        L()
        L("--- Activity & friction ---")
        L(f"  ALPHA (numerical)  = {_fmt(p.alpha)}")
        L(f"  MU (linear friction) = {_fmt(p.mu)}")
        L()
        L("#" * 58)
        self.f.flush()

    def log(self, diag: dict, time_step: int, num_disclinations: int = 0) -> bool:
        """One SimIO::Log line (src/sim_io.cc:78-165) from lbm_diagnostics."""
        mass, px, py, pz = diag["mass"], diag["px"], diag["py"], diag["pz"]
        ke, fe = diag["kinetic"], diag["free_energy"]
        te = ke + fe
        if diag["u2"] != 0.0:
            rel = diag["du2"] / diag["u2"]
        else:   # IEEE x / 0: inf, and 0 / 0 gives the x86 default NaN (sign bit set)
            rel = math.inf if diag["du2"] > 0.0 else -math.nan
        self._line(f"Time {time_step}: Mass: {_fmt(mass)}, Px: {_fmt(px)}, Py: {_fmt(py)}, Pz: {_fmt(pz)}, "
                   f"Kinetic Energy: {_fmt(ke)}, Total Energy: {_fmt(te)}, Relative Error: {_fmt(rel)}, "
                   f"NumDisclinations: {num_disclinations}, Nematic Energy: {_fmt(te - ke)}, "
                   f"Max |u|: {_fmt(diag['umax'])}")
        self.f.flush()
        diverged = any(math.isnan(v) for v in (mass, px, py, pz))
        if diverged:
            self._line(f"DIVERGED at time step {time_step} — aborting.")
            self.f.flush()
            return False
        return True

    def close(self):
        if self.f:
            self._line("LBM program exited.")
            self.f.close()
            self.f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def run(cfg: abi.LbmConfig, q0, steps: int, save_interval: int, bc_name: str, log_path="lbm.log",
        rho0=None, u0=None, on_save=None) -> int:
    """The reference's main loop (src/main.cc:9-31) over a device solver:
    every save_interval steps the state is logged (and handed to on_save,
    the place of Export), then one full step. Returns the number of steps
    completed; stops at the first Log that reports divergence."""
    from .solver import DeviceSolver, device_info
    with DeviceSolver(cfg) as s, SimLog(log_path) as log:
        log.setup_summary(cfg, bc_name, device_info())
        s.upload_state(rho=rho0, u=u0, q=q0)   # the previous-velocity reference starts at zero, as AnalysisFields
        for t in range(steps):
            if t % save_interval == 0:
                if on_save is not None:
                    on_save(t, s)
                if not log.log(s.diagnostics(), t):
                    return t
            s.step(1)
        return steps
