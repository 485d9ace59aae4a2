"""Host interface of the 3D annulus step (rot3dmg.f, and rot3dmghor.f with
model=1) on the HIP library: one object per run, methods named after the reference's routines
(Avorticity, ..., poisson1/2/3, update), fields as numpy arrays of shape
(np+1, nz+2, nr+2) (X[k, j, i] is X(i, j, k) of rot3dmg.f).

The compute path is libvvmi.so (include/vvmi3.h); there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib

FIELDS = ["An", "A", "Ap", "Bn", "B", "Bp", "Cn", "C", "Cp", "thn", "th", "thp",
          "uel", "vel", "wel", "uelp", "velp", "welp", "urh", "vrh", "wrh"]
FIELD_IDS = {n: q for q, n in enumerate(FIELDS)}


class VV3Config(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nz", ctypes.c_int), ("np", ctypes.c_int),
                ("dt", ctypes.c_double), ("r0", ctypes.c_double), ("rmax", ctypes.c_double),
                ("zmax", ctypes.c_double), ("Ra", ctypes.c_double), ("Re", ctypes.c_double),
                ("Ptl", ctypes.c_double), ("Gr", ctypes.c_double), ("dr", ctypes.c_double),
                ("dz", ctypes.c_double), ("dp", ctypes.c_double), ("model", ctypes.c_int)]

MODEL_ROT3DMG = 0   # rot3dmg.f: rotating annulus
MODEL_HOR = 1       # rot3dmghor.f: stationary horizontal annulus


@dataclass
class Annulus3DConfig:
    """The scalars of rot3dmg.f's input deck (rot3dmg.f:59-86); model 1
    steps rot3dmghor.f's physics instead (Re unused)."""
    nr: int
    nz: int
    np: int
    dt: float
    Ra: float
    Re: float
    Ptl: float
    Gr: float
    r0: float = 0.5
    rmax: float = 1.0
    zmax: float = 0.5
    model: int = MODEL_ROT3DMG

    def c(self) -> VV3Config:
        return VV3Config(self.nr, self.nz, self.np, self.dt, self.r0, self.rmax, self.zmax,
                         self.Ra, self.Re, self.Ptl, self.Gr, 0.0, 0.0, 0.0, self.model)


_P = ctypes.POINTER(ctypes.c_double)


def _bind():
    L = lib()
    if getattr(L, "_vv3_bound", False):
        return L
    vp = ctypes.c_void_p
    L.vv3_create.argtypes = [ctypes.POINTER(VV3Config), ctypes.c_int, ctypes.POINTER(vp)]
    L.vv3_destroy.argtypes = [vp]
    L.vv3_put.argtypes = [vp, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]
    L.vv3_get.argtypes = [vp, ctypes.c_int, _P, ctypes.c_int, ctypes.c_int]
    L.vv3_vorticity.argtypes = [vp, ctypes.c_int]
    L.vv3_temperature.argtypes = [vp]
    L.vv3_heat.argtypes = [vp, _P]
    L.vv3_fluxes.argtypes = [vp, _P, _P]
    L.vv3_energy.argtypes = [vp, _P]
    L.vv3_local_nu.argtypes = [vp, _P, _P]
    L.vv3_modes.argtypes = [vp, ctypes.c_int, _P, _P]
    L.vv3_rhs.argtypes = [vp, ctypes.c_int]
    L.vv3_poisson.argtypes = [vp, ctypes.c_int]
    L.vv3_update.argtypes = [vp]
    L.vv3_step.argtypes = [vp, ctypes.c_int, _P]
    L._vv3_bound = True
    return L


class Annulus3D:
    """rot3dmg.f's state on one GPU.  The routines act on the device fields;
    put/get move whole fields."""

    def __init__(self, cfg: Annulus3DConfig, device: int = 0):
        self.cfg = cfg
        self.shape = (cfg.np + 1, cfg.nz + 2, cfg.nr + 2)
        self._L = _bind()
        self._h = ctypes.c_void_p()
        c = cfg.c()
        check(self._L.vv3_create(ctypes.byref(c), device, ctypes.byref(self._h)), "vv3_create")

    def close(self):
        if self._h:
            self._L.vv3_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def put(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"{name}: shape {a.shape}, expected {self.shape}")
        check(self._L.vv3_put(self._h, FIELD_IDS[name], a.ctypes.data_as(_P), 0, 0), "vv3_put")

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.shape)
        check(self._L.vv3_get(self._h, FIELD_IDS[name], out.ctypes.data_as(_P), 0, 0), "vv3_get")
        return out

    # rot3dmg.f routine names
    def Avorticity(self):
        check(self._L.vv3_vorticity(self._h, 1), "vv3_vorticity")

    def Bvorticity(self):
        check(self._L.vv3_vorticity(self._h, 2), "vv3_vorticity")

    def Cvorticity(self):
        check(self._L.vv3_vorticity(self._h, 3), "vv3_vorticity")

    def vorticity(self):
        """Avorticity, Bvorticity and Cvorticity in one launch."""
        check(self._L.vv3_vorticity(self._h, 0), "vv3_vorticity")

    def temperature(self):
        check(self._L.vv3_temperature(self._h), "vv3_temperature")

    def heat(self) -> float:
        t = ctypes.c_double()
        check(self._L.vv3_heat(self._h, ctypes.byref(t)), "vv3_heat")
        return t.value

    def fluxes(self) -> tuple[float, float]:
        """(trans1, trans2) of the last heat evaluation: rot3dmghor.f's
        heat(trans1,trans2) (trans2 = 0 for rot3dmg.f)."""
        t1, t2 = ctypes.c_double(), ctypes.c_double()
        check(self._L.vv3_fluxes(self._h, ctypes.byref(t1), ctypes.byref(t2)), "vv3_fluxes")
        return t1.value, t2.value

    def energy(self) -> float:
        """Kinetic energy of the current velocities (vv3_energy)."""
        t = ctypes.c_double()
        check(self._L.vv3_energy(self._h, ctypes.byref(t)), "vv3_energy")
        return t.value

    def local_nu(self) -> tuple[np.ndarray, np.ndarray]:
        """Local Nusselt numbers at mid-height of the inner and outer
        cylinder, one per plane k = 0..np (vv3_local_nu)."""
        a, b = np.empty(self.cfg.np + 1), np.empty(self.cfg.np + 1)
        check(self._L.vv3_local_nu(self._h, a.ctypes.data_as(_P), b.ctypes.data_as(_P)), "vv3_local_nu")
        return a, b

    def modes(self, nmax: int = 6) -> tuple[np.ndarray, np.ndarray]:
        """Fourier-mode volume integrals (cos, sin) of the temperature for
        n = 1..nmax (vv3_modes)."""
        c, s = np.empty(nmax), np.empty(nmax)
        check(self._L.vv3_modes(self._h, nmax, c.ctypes.data_as(_P), s.ctypes.data_as(_P)), "vv3_modes")
        return c, s

    def rhsw(self):
        check(self._L.vv3_rhs(self._h, 1), "vv3_rhs")

    def rhsv(self):
        check(self._L.vv3_rhs(self._h, 2), "vv3_rhs")

    def rhsu(self):
        check(self._L.vv3_rhs(self._h, 3), "vv3_rhs")

    def poisson1(self):
        check(self._L.vv3_poisson(self._h, 1), "vv3_poisson")

    def poisson2(self):
        check(self._L.vv3_poisson(self._h, 2), "vv3_poisson")

    def poisson3(self):
        check(self._L.vv3_poisson(self._h, 3), "vv3_poisson")

    def update(self):
        check(self._L.vv3_update(self._h), "vv3_update")

    def step(self, n: int = 1) -> float:
        """n passes of the reference time loop (rot3dmg.f:240-253); returns
        trans1 of the last pass."""
        t = ctypes.c_double()
        check(self._L.vv3_step(self._h, n, ctypes.byref(t)), "vv3_step")
        return t.value
