"""ctypes wrapper of the 3D oracle (vv3_oracle.c in liboracle.so).

Test infrastructure only: the tests, the smoke check and the benchmark's
CPU leg use it; no product code imports it.  Fields are numpy arrays of
shape (np+1, ldj, ldi) (C order: X[k, j, i] is X(i, j, k) of rot3dmg.f).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FIELDS = ["An", "A", "Ap", "Bn", "B", "Bp", "Cn", "C", "Cp", "thn", "th", "thp",
          "uel", "vel", "wel", "uelp", "velp", "welp", "urh", "vrh", "wrh"]


class V3Params(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nz", ctypes.c_int), ("np", ctypes.c_int), ("ldi", ctypes.c_int),
                ("ldj", ctypes.c_int), ("dt", ctypes.c_double), ("dr", ctypes.c_double),
                ("dz", ctypes.c_double), ("dp", ctypes.c_double), ("Ra", ctypes.c_double),
                ("Re", ctypes.c_double), ("Ptl", ctypes.c_double), ("Gr", ctypes.c_double),
                ("r0", ctypes.c_double), ("rmax", ctypes.c_double), ("model", ctypes.c_int)]


class V3Fields(ctypes.Structure):
    _fields_ = [(n, ctypes.POINTER(ctypes.c_double)) for n in FIELDS]


def lib():
    global _LIB
    if _LIB is None:
        _LIB = ctypes.CDLL(os.path.join(HERE, "liboracle.so"))
        for fn in ("v3_avorticity", "v3_bvorticity", "v3_cvorticity", "v3_temperature", "v3_update",
                   "v3_rhsw", "v3_rhsv", "v3_rhsu"):
            getattr(_LIB, fn).restype = None
        _LIB.v3_heat.restype = ctypes.c_double
        _LIB.v3_heat2.restype = None
        _LIB.v3_energy.restype = ctypes.c_double
        _LIB.v3_local_nu.restype = None
        _LIB.v3_modes.restype = None
        _LIB.v3_residual.restype = ctypes.c_double
        _LIB.v3_poisson_exact.restype = ctypes.c_int
    return _LIB


def _P(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Run3:
    """One 3D state: parameters and the 21 fields of rot3dmg.f's COMMON
    blocks (plus the three Poisson right-hand sides).  model 0: rot3dmg.f;
    1: rot3dmghor.f (stationary horizontal annulus)."""

    def __init__(self, nr, nz, np_, dt, Ra, Re, Ptl, Gr, r0=0.5, rmax=1.0, zmax=0.5, ldi=None, ldj=None,
                 model=0):
        ldi = ldi or nr + 2
        ldj = ldj or nz + 2
        pi = np.arccos(-1.0)
        self.p = V3Params(nr, nz, np_, ldi, ldj, dt, (rmax - r0) / nr, zmax / nz, 2.0 * pi / np_,
                          Ra, Re, Ptl, Gr, r0, rmax, model)
        self.shape = (np_ + 1, ldj, ldi)
        self.f = {n: np.zeros(self.shape) for n in FIELDS}
        self._sync()

    def _sync(self):
        self.c = V3Fields(*[_P(self.f[n]) for n in FIELDS])

    def __getitem__(self, name):
        return self.f[name]

    def __setitem__(self, name, val):
        self.f[name][...] = val

    def call(self, name):
        getattr(lib(), "v3_" + name)(ctypes.byref(self.p), ctypes.byref(self.c))

    def heat(self):
        return lib().v3_heat(ctypes.byref(self.p), ctypes.byref(self.c))

    def heat2(self):
        """(trans1, trans2) of the model's heat routine (trans2 = 0 for
        rot3dmg.f)."""
        out = np.zeros(2)
        lib().v3_heat2(ctypes.byref(self.p), ctypes.byref(self.c), _P(out))
        return float(out[0]), float(out[1])

    def energy(self):
        """Kinetic energy of the velocities (the build's definition)."""
        return lib().v3_energy(ctypes.byref(self.p), ctypes.byref(self.c))

    def local_nu(self):
        """(inner, outer) local Nusselt numbers at mid-height, k = 0..np."""
        a, b = np.zeros(self.p.np + 1), np.zeros(self.p.np + 1)
        lib().v3_local_nu(ctypes.byref(self.p), ctypes.byref(self.c), _P(a), _P(b))
        return a, b

    def modes(self, nmax=6):
        """(c, s): Fourier-mode integrals n = 1..nmax of the temperature."""
        c, s = np.zeros(nmax), np.zeros(nmax)
        lib().v3_modes(ctypes.byref(self.p), ctypes.byref(self.c), nmax, _P(c), _P(s))
        return c, s

    def residual(self, which, psi, rhs):
        res = np.zeros(self.shape)
        m = lib().v3_residual(ctypes.byref(self.p), which, _P(np.ascontiguousarray(psi)),
                              _P(np.ascontiguousarray(rhs)), _P(res))
        return m, res

    def poisson_exact(self, which, rhs):
        psi = np.zeros(self.shape)
        st = lib().v3_poisson_exact(ctypes.byref(self.p), which, _P(psi), _P(np.ascontiguousarray(rhs)))
        if st:
            raise RuntimeError("v3_poisson_exact: singular mode system")
        return psi
