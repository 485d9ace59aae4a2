"""ctypes wrapper of the sector oracle (vvs_oracle.c in liboracle.so):
sect360v2mg.f's rotating annulus with a radial barrier.

Test infrastructure only: the tests, the smoke check and the benchmark's
CPU leg use it; no product code imports it.  Fields are numpy arrays of
shape (np+2, ldj, ldi) (C order: X[k, j, i] is X(i, j, k)); the planes
k = 0 and k = np+1 lie on the barrier's two faces.
"""
from __future__ import annotations

import ctypes

import numpy as np

from oracle3 import FIELDS, V3Fields, V3Params, _P, lib as _lib3

_BOUND = False


def lib():
    global _BOUND
    L = _lib3()
    if not _BOUND:
        for fn in ("vs_avorticity", "vs_bvorticity", "vs_cvorticity", "vs_temperature", "vs_update",
                   "vs_rhsw", "vs_rhsv", "vs_rhsu", "vs_ranges"):
            getattr(L, fn).restype = None
        L.vs_heat.restype = ctypes.c_double
        L.vs_residual.restype = ctypes.c_double
        L.vs_poisson_exact.restype = ctypes.c_int
        L.vs_poisson.restype = ctypes.c_int
        _BOUND = True
    return L


class Sector:
    """One state of sect360v2mg.f: the parameters (rot3dmg.f's deck) and the
    21 fields of its COMMON blocks plus the three right-hand sides."""

    def __init__(self, nr, nz, np_, dt, Ra, Re, Ptl, Gr, r0=0.5, rmax=1.0, zmax=0.5, ldi=None, ldj=None):
        ldi = ldi or nr + 2
        ldj = ldj or nz + 2
        pi = np.arccos(-1.0)
        self.p = V3Params(nr, nz, np_, ldi, ldj, dt, (rmax - r0) / nr, zmax / nz, 2.0 * pi / np_,
                          Ra, Re, Ptl, Gr, r0, rmax, 0)
        self.shape = (np_ + 2, ldj, ldi)
        self.f = {n: np.zeros(self.shape) for n in FIELDS}
        self.c = V3Fields(*[_P(self.f[n]) for n in FIELDS])

    def __getitem__(self, name):
        return self.f[name]

    def __setitem__(self, name, val):
        self.f[name][...] = val

    def call(self, name):
        getattr(lib(), "vs_" + name)(ctypes.byref(self.p), ctypes.byref(self.c))

    def heat(self):
        return lib().vs_heat(ctypes.byref(self.p), ctypes.byref(self.c))

    def ranges(self, which):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().vs_ranges(ctypes.byref(self.p), which, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def residual(self, which, psi, rhs):
        res = np.zeros(self.shape)
        m = lib().vs_residual(ctypes.byref(self.p), which, _P(np.ascontiguousarray(psi)),
                              _P(np.ascontiguousarray(rhs)), _P(res))
        return m, res

    def poisson_exact(self, which, rhs):
        x = np.zeros(self.shape)
        if lib().vs_poisson_exact(ctypes.byref(self.p), which, _P(x), _P(np.ascontiguousarray(rhs))):
            raise RuntimeError("vs_poisson_exact: singular system")
        return x

    def poisson(self, which):
        psi, psip, rhs = PSI[which]
        if lib().vs_poisson(ctypes.byref(self.p), which, _P(self.f[psi]), _P(self.f[psip]), _P(self.f[rhs])):
            raise RuntimeError("vs_poisson: singular system")

    def step(self):
        """One pass of the time loop (sect360v2mg.f:235-249); returns trans1."""
        for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
            self.call(rt)
        t = self.heat()
        self.call("rhsw")
        self.poisson(1)
        self.call("rhsv")
        self.poisson(2)
        self.call("rhsu")
        self.poisson(3)
        self.call("update")
        return t


PSI = {1: ("wel", "welp", "wrh"), 2: ("vel", "velp", "vrh"), 3: ("uel", "uelp", "urh")}
