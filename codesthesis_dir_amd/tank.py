"""Host interface of the 3D horizontal-convection tank step on the HIP
library (include/vvtank.h): one object per run, methods named after the
routines of the step (derivatives, vorticity, temperature, heat, rhs,
poisson, update), fields as numpy arrays of shape (np+2, nz+2, nr+2)
(X[k, j, i] is X(i, j, k)).

The compute path is libvvmi.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib
from .velvor3 import FIELD_IDS, FIELDS  # noqa: F401  (the tank uses rot3dmg.f's fields)

_P = ctypes.POINTER(ctypes.c_double)
KERNELS = ["deriv_yz", "deriv_x", "avort", "bvort", "cvort", "temp", "heat", "rhs", "psi_pre", "yfwd", "zfwd",
           "xsolve", "zinv", "yinv", "update", "yfft_fwd", "yfft_inv", "explicit"]
# the work of these kernels is counted in floating-point operations, of the others in bytes
FLOP_KERNELS = ("yfwd", "zfwd", "zinv", "yinv")


class VVTConfig(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nz", ctypes.c_int), ("np", ctypes.c_int), ("dt", ctypes.c_double),
                ("lx", ctypes.c_double), ("lz", ctypes.c_double), ("ly", ctypes.c_double),
                ("Re", ctypes.c_double), ("Ptl", ctypes.c_double), ("Gr", ctypes.c_double)]


@dataclass
class TankConfig:
    """A tank of nr x nz x np intervals (x across, z vertical, y along),
    sizes lx x lz x ly; Re = sqrt(Ta), Gr Ptl = Ra."""
    nr: int
    nz: int
    np: int
    dt: float
    Re: float
    Ptl: float
    Gr: float
    lx: float = 1.0
    lz: float = 1.0
    ly: float = 4.0

    def c(self) -> VVTConfig:
        return VVTConfig(self.nr, self.nz, self.np, self.dt, self.lx, self.lz, self.ly, self.Re, self.Ptl,
                         self.Gr)


def _bind():
    L = lib()
    if getattr(L, "_vvt_bound", False):
        return L
    vp = ctypes.c_void_p
    L.vvt_create.argtypes = [ctypes.POINTER(VVTConfig), ctypes.c_int, ctypes.POINTER(vp)]
    for fn in ("vvt_destroy", "vvt_derivatives", "vvt_vorticity", "vvt_temperature", "vvt_update"):
        getattr(L, fn).argtypes = [vp]
    L.vvt_put.argtypes = [vp, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]
    L.vvt_get.argtypes = [vp, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]
    L.vvt_get_derivative.argtypes = [vp, ctypes.c_int, _P]
    L.vvt_heat.argtypes = [vp, _P, _P]
    L.vvt_rhs.argtypes = [vp, ctypes.c_int]
    L.vvt_poisson.argtypes = [vp, ctypes.c_int]
    L.vvt_step.argtypes = [vp, ctypes.c_int, _P, _P]
    L.vvt_energy.argtypes = [vp, _P]
    L.vvt_profile.argtypes = [vp, ctypes.c_int]
    L.vvt_profile_read.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _P, _P]
    L._vvt_bound = True
    return L


class Tank:
    """One tank run resident on a GPU."""

    def __init__(self, cfg: TankConfig, device: int = 0):
        self.cfg = cfg
        self._lib = _bind()
        h = ctypes.c_void_p()
        c = cfg.c()
        check(self._lib.vvt_create(ctypes.byref(c), device, ctypes.byref(h)), "vvt_create")
        self._h = h
        self.shape = (cfg.np + 2, cfg.nz + 2, cfg.nr + 2)

    def close(self):
        if self._h:
            self._lib.vvt_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def put(self, name: str, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"{name}: shape {a.shape}, expected {self.shape}")
        check(self._lib.vvt_put(self._h, FIELD_IDS[name], a.ctypes.data_as(_P), 0, 0), "vvt_put")

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.shape)
        check(self._lib.vvt_get(self._h, FIELD_IDS[name], out.ctypes.data_as(_P), 0, 0), "vvt_get")
        return out

    def derivative(self, d: int) -> np.ndarray:
        out = np.empty(self.shape)
        check(self._lib.vvt_get_derivative(self._h, d, out.ctypes.data_as(_P)), "vvt_get_derivative")
        return out

    def derivatives(self):
        check(self._lib.vvt_derivatives(self._h), "vvt_derivatives")

    def vorticity(self):
        check(self._lib.vvt_vorticity(self._h), "vvt_vorticity")

    def temperature(self):
        check(self._lib.vvt_temperature(self._h), "vvt_temperature")

    def heat(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(self._lib.vvt_heat(self._h, ctypes.byref(a), ctypes.byref(b)), "vvt_heat")
        return a.value, b.value

    def rhs(self, which: int):
        check(self._lib.vvt_rhs(self._h, which), "vvt_rhs")

    def poisson(self, which: int):
        check(self._lib.vvt_poisson(self._h, which), "vvt_poisson")

    def update(self):
        check(self._lib.vvt_update(self._h), "vvt_update")

    def step(self, n: int = 1):
        a, b = ctypes.c_double(), ctypes.c_double()
        check(self._lib.vvt_step(self._h, n, ctypes.byref(a), ctypes.byref(b)), "vvt_step")
        return a.value, b.value

    def energy(self) -> float:
        e = ctypes.c_double()
        check(self._lib.vvt_energy(self._h, ctypes.byref(e)), "vvt_energy")
        return e.value

    def profile(self, enable: bool):
        check(self._lib.vvt_profile(self._h, int(enable)), "vvt_profile")

    def profile_read(self):
        out = {}
        for k, name in enumerate(KERNELS):
            n, ms, b = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
            check(self._lib.vvt_profile_read(self._h, k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)),
                  "vvt_profile_read")
            unit = "flops" if name in FLOP_KERNELS else "bytes"
            out[name] = {"launches": n.value, "ms": ms.value, unit: b.value}
        return out
