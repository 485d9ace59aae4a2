"""Host-side mirror of the reference's periodic-box routines over libvvmi.

`PeriodicBox` holds one run on the GPU and exposes the reference's time-loop
routines under their own names (velvor2dperiodicrbconv.f):

    Cvorticity1 / Cvorticity   -> cvorticity(first)      rb.f:584 / rb.f:786
    temperature1 / temperature -> temperature(first)     rb.f:1056 / rb.f:1220
    heat(trans1, trans2)       -> heat()                 rb.f:1452
    rhsv(vrh), rhsu(urh)       -> rhs()                  rb.f:1629 / rb.f:1679
    poisson2 / poisson3        -> poisson(2) / poisson(3) rb.f:1796 / rb.f:1930
    update                     -> update()               rb.f:1557
    the time-loop body         -> step(n)                rb.f:341-470

Fields move between host and device as numpy arrays of shape (np+1, ld) in C
order, i.e. the reference's column-major F(0:ld-1,1,0:np).  Every call goes
to the HIP library; there is no CPU path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class BoxConfig:
    nr: int = 64
    np: int = 128
    dt: float = 1e-5
    r0: float = 0.5
    rmax: float = 1.0
    width: float = 1.0
    Ra: float = 1e5
    Pr: float = 1.0
    robert: float = 0.02
    rtol: float = 1e-10
    max_cycles: int = 60
    nu1: int = 2
    nu2: int = 2
    stats: int = 1
    geom: int = 0          # 0: periodic box (rb.f); 1: rotating annulus plane (ann.f)
    Re: float = 0.0        # rotational Reynolds number (annulus)
    direct_cols: int = 0  # direct level at <= this many columns (0: direct for np <= 8192, else 128)
    precision: int = 0    # 0: FP64 V-cycle; 1: FP64 defect correction around an FP32 V-cycle
    lines: int = 0        # 0: the reference's line recurrences (bit-exact); 1: parallel scan (to rounding)

    def c(self) -> L.VVConfig:
        return L.VVConfig(self.nr, self.np, self.dt, self.r0, self.rmax, self.width, self.Ra,
                          self.Pr, self.robert, self.rtol, self.max_cycles, self.nu1, self.nu2,
                          self.stats, self.geom, self.Re, self.direct_cols, self.precision, self.lines)


def AnnulusConfig(**kw) -> BoxConfig:
    """Configuration of the rotating-annulus plane of velvor2dnewsch3.4.f:
    dp = 2 pi/np, Robert coefficient 0.05 (ann.f:1055)."""
    kw.setdefault("robert", 0.05)
    return BoxConfig(geom=1, **kw)


def _P(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class PeriodicBox:
    """One 2D periodic Rayleigh-Benard run resident on a GPU."""

    def __init__(self, cfg: BoxConfig, device: int = 0):
        self.cfg = cfg
        self._c = cfg.c()
        self._run = ctypes.c_void_p()
        nlev = ctypes.c_int(0)
        L.check(L.lib().vv_create(ctypes.byref(self._c), device, ctypes.byref(self._run),
                                  ctypes.byref(nlev)), "vv_create")
        self.nlev = nlev.value
        self.ld = cfg.nr + 2
        self.steps = 0

    def close(self):
        if self._run:
            L.lib().vv_destroy(self._run)
            self._run = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- fields
    def put(self, name: str, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape[0] != self.cfg.np + 1 or a.shape[1] < self.cfg.nr + 2:
            raise ValueError(f"field {name}: shape {a.shape} does not hold (np+1, >= nr+2)")
        L.check(L.lib().vv_put(self._run, L.FIELD_IDS[name], _P(a), a.shape[1]), "vv_put")

    def get(self, name: str) -> np.ndarray:
        a = np.zeros((self.cfg.np + 1, self.ld))
        L.check(L.lib().vv_get(self._run, L.FIELD_IDS[name], _P(a), self.ld), "vv_get")
        return a

    def start(self):
        """Previous time levels copy the current ones (rb.f:315-329)."""
        L.check(L.lib().vv_start(self._run), "vv_start")
        self.steps = 0

    # ------------------------------------------------- reference routines
    def cvorticity(self, first: bool = False):
        L.check(L.lib().vv_cvorticity(self._run, int(first)), "vv_cvorticity")

    def temperature(self, first: bool = False):
        L.check(L.lib().vv_temperature(self._run, int(first)), "vv_temperature")

    def heat(self):
        t1, t2 = ctypes.c_double(), ctypes.c_double()
        L.check(L.lib().vv_heat(self._run, ctypes.byref(t1), ctypes.byref(t2)), "vv_heat")
        return t1.value, t2.value

    def energy(self) -> float:
        """Kinetic energy of the current velocities (vv_energy)."""
        ke = ctypes.c_double()
        L.check(L.lib().vv_energy(self._run, ctypes.byref(ke)), "vv_energy")
        return ke.value

    def rhs(self):
        L.check(L.lib().vv_rhs(self._run), "vv_rhs")

    def poisson(self, which: int):
        cyc = (ctypes.c_int * 2)()
        rel = (ctypes.c_double * 2)()
        st = L.check(L.lib().vv_poisson(self._run, which, cyc, rel), "vv_poisson")
        return st, list(cyc), list(rel)

    def poisson_host(self, which: int, psi: np.ndarray, psip: np.ndarray, rhs: np.ndarray):
        """poisson2/poisson3 with the reference argument list on host arrays
        (updated in place)."""
        for a in (psi, psip, rhs):
            if a.dtype != np.float64 or not a.flags.c_contiguous:
                raise ValueError("arrays must be C-contiguous float64")
        cyc = (ctypes.c_int * 2)()
        rel = (ctypes.c_double * 2)()
        st = L.check(L.lib().vv_poisson_host(self._run, which, _P(psi), _P(psip), _P(rhs),
                                             psi.shape[1], cyc, rel), "vv_poisson_host")
        return st, cyc[0], rel[0]

    def update(self) -> int:
        n = ctypes.c_int(0)
        L.check(L.lib().vv_update(self._run, ctypes.byref(n)), "vv_update")
        self.steps += 1
        return n.value

    def step(self, n: int = 1) -> L.VVStepInfo:
        info = L.VVStepInfo()
        st = L.check(L.lib().vv_step(self._run, n, ctypes.byref(info)), "vv_step")
        self.steps += n
        info.status = st
        return info

    def reset(self, cfg: "BoxConfig"):
        """Re-arm the run for cfg (same grid, geometry and solver; new
        physics): fields and solver state as after creation, tables kept."""
        c = cfg.c()
        L.check(L.lib().vv_reset(self._run, ctypes.byref(c)), "vv_reset")
        self.cfg = cfg
        self.steps = 0

    def solve_history(self, max_cycles: int = 64):
        """Relative residual after each cycle of the last solve of v and of
        u: (v history, u history)."""
        v, u = np.zeros(max_cycles), np.zeros(max_cycles)
        n = (ctypes.c_int * 2)()
        L.check(L.lib().vv_solve_history(self._run, _P(v), _P(u), max_cycles, n), "vv_solve_history")
        return v[:min(n[0], max_cycles)], u[:min(n[1], max_cycles)]

    def heat_history(self, n: int) -> np.ndarray:
        out = np.zeros((n, 2))
        L.check(L.lib().vv_heat_history(self._run, _P(out), n), "vv_heat_history")
        return out

    def last_step_ms(self) -> float:
        ms = ctypes.c_double()
        L.check(L.lib().vv_last_step_ms(self._run, ctypes.byref(ms)), "vv_last_step_ms")
        return ms.value

    KERNELS = ["compact", "df", "heat", "rhs", "zebra", "restrict", "prolong", "norm",
               "direct", "out", "update"]

    def profile(self, enable: bool):
        """Start (clearing totals) or stop per-kernel HIP event timing."""
        L.check(L.lib().vv_profile(self._run, int(enable)), "vv_profile")

    def profile_read(self) -> dict:
        out = {}
        for kid, name in enumerate(self.KERNELS):
            n = ctypes.c_longlong()
            ms = ctypes.c_double()
            b = ctypes.c_double()
            L.check(L.lib().vv_profile_read(self._run, kid, ctypes.byref(n), ctypes.byref(ms),
                                            ctypes.byref(b)), "vv_profile_read")
            out[name] = {"launches": n.value, "ms": ms.value, "bytes": b.value}
        return out

    def nusselt(self, trans):
        """Nu = trans*(rmax-r0)/width in the box (rb.f:386-388);
        -r0 log(r0/rmax) trans/(2 pi) in the annulus (ann.f:376-377)."""
        if self.cfg.geom == 1:
            c = self.cfg
            return -c.r0 * np.log(c.r0 / c.rmax) * np.asarray(trans) / (2.0 * np.pi)
        return np.asarray(trans) * (self.cfg.rmax - self.cfg.r0) / self.cfg.width
