"""Host interface of the barrier-sector step of sect360v2mg.f on the HIP
library (include/vvsector.h): one object per run, methods named after the
reference's routines (vorticity, temperature, heat, rhs, poisson, update),
fields as numpy arrays of shape (np+2, nz+2, nr+2) (X[k, j, i] is X(i, j, k);
the planes k = 0 and np+1 lie on the barrier's faces).

The compute path is libvvmi.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib
from .velvor3 import FIELD_IDS, VV3Config

_P = ctypes.POINTER(ctypes.c_double)


@dataclass
class SectorConfig:
    """sect360v2mg.f's input deck (sect360v2mg.f:53-82): nr x nz x np
    intervals, radii r0 < rmax, height zmax; Re = sqrt(Ta)."""
    nr: int
    nz: int
    np: int
    dt: float
    Ra: float
    Re: float
    Ptl: float
    Gr: float = 0.0
    r0: float = 0.5
    rmax: float = 1.0
    zmax: float = 0.5

    def c(self) -> VV3Config:
        return VV3Config(self.nr, self.nz, self.np, self.dt, self.r0, self.rmax, self.zmax, self.Ra, self.Re,
                         self.Ptl, self.Gr, 0.0, 0.0, 0.0, 0)


def _bind():
    L = lib()
    if getattr(L, "_vvs_bound", False):
        return L
    vp = ctypes.c_void_p
    L.vvs_create.argtypes = [ctypes.POINTER(VV3Config), ctypes.c_int, ctypes.POINTER(vp)]
    for fn in ("vvs_destroy", "vvs_temperature", "vvs_update"):
        getattr(L, fn).argtypes = [vp]
    L.vvs_put.argtypes = [vp, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]
    L.vvs_get.argtypes = [vp, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]
    for fn in ("vvs_vorticity", "vvs_rhs", "vvs_poisson"):
        getattr(L, fn).argtypes = [vp, ctypes.c_int]
    L.vvs_heat.argtypes = [vp, _P]
    L.vvs_local_nu.argtypes = [vp, _P, _P]
    L.vvs_step.argtypes = [vp, ctypes.c_int, _P]
    L._vvs_bound = True
    return L


class Sector:
    """One barrier-sector run resident on a GPU."""

    def __init__(self, cfg: SectorConfig, device: int = 0):
        self.cfg = cfg
        self._lib = _bind()
        h = ctypes.c_void_p()
        c = cfg.c()
        check(self._lib.vvs_create(ctypes.byref(c), device, ctypes.byref(h)), "vvs_create")
        self._h = h
        self.shape = (cfg.np + 2, cfg.nz + 2, cfg.nr + 2)

    def close(self):
        if self._h:
            self._lib.vvs_destroy(self._h)
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
        check(self._lib.vvs_put(self._h, FIELD_IDS[name], a.ctypes.data_as(_P), 0, 0), "vvs_put")

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.shape)
        check(self._lib.vvs_get(self._h, FIELD_IDS[name], out.ctypes.data_as(_P), 0, 0), "vvs_get")
        return out

    def vorticity(self, which: int = 0):
        check(self._lib.vvs_vorticity(self._h, which), "vvs_vorticity")

    def temperature(self):
        check(self._lib.vvs_temperature(self._h), "vvs_temperature")

    def heat(self) -> float:
        t = ctypes.c_double()
        check(self._lib.vvs_heat(self._h, ctypes.byref(t)), "vvs_heat")
        return t.value

    def local_nu(self):
        """(inner, outer) local Nusselt numbers at mid-height, k = 0..np."""
        a, b = np.empty(self.cfg.np + 1), np.empty(self.cfg.np + 1)
        check(self._lib.vvs_local_nu(self._h, a.ctypes.data_as(_P), b.ctypes.data_as(_P)), "vvs_local_nu")
        return a, b

    def rhs(self, which: int):
        check(self._lib.vvs_rhs(self._h, which), "vvs_rhs")

    def poisson(self, which: int):
        check(self._lib.vvs_poisson(self._h, which), "vvs_poisson")

    def update(self):
        check(self._lib.vvs_update(self._h), "vvs_update")

    def step(self, n: int = 1) -> float:
        t = ctypes.c_double()
        check(self._lib.vvs_step(self._h, n, ctypes.byref(t)), "vvs_step")
        return t.value
