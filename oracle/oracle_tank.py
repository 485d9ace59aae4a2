"""ctypes wrapper of the 3D tank oracle (vvt_oracle.c in liboracle.so).

Test infrastructure only: the tests, the smoke check and the benchmark's CPU
leg use it; no product code imports it.  Fields are numpy arrays of shape
(np+2, ldj, ldi) in C order: X[k, j, i] is X(i, j, k), k = 0..np+1.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FIELDS = ["An", "A", "Ap", "Bn", "B", "Bp", "Cn", "C", "Cp", "thn", "th", "thp",
          "uel", "vel", "wel", "uelp", "velp", "welp", "urh", "vrh", "wrh"]
PSI = {1: ("wel", "welp", "wrh"), 2: ("vel", "velp", "vrh"), 3: ("uel", "uelp", "urh")}
_D = ctypes.POINTER(ctypes.c_double)


class VtParams(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nz", ctypes.c_int), ("np", ctypes.c_int), ("ldi", ctypes.c_int),
                ("ldj", ctypes.c_int), ("dt", ctypes.c_double), ("dr", ctypes.c_double),
                ("dz", ctypes.c_double), ("dp", ctypes.c_double), ("Re", ctypes.c_double),
                ("Ptl", ctypes.c_double), ("Gr", ctypes.c_double)]


class VtFields(ctypes.Structure):
    _fields_ = [(n, _D) for n in FIELDS]


class VtDerivs(ctypes.Structure):
    _fields_ = [("d", _D * 12)]


def lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(os.path.join(HERE, "liboracle.so"))
        for fn in ("vt_avorticity", "vt_bvorticity", "vt_cvorticity", "vt_temperature", "vt_update", "vt_rhsw",
                   "vt_rhsv", "vt_rhsu", "vt_derivatives", "vt_heat", "vt_apply", "vt_psi_bounds",
                   "vt_compact_line", "vt_compact_tables", "vt_ranges"):
            getattr(L, fn).restype = None
        L.vt_residual.restype = ctypes.c_double
        L.vt_energy.restype = ctypes.c_double
        L.vt_compact_line.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _D, _D]
        _LIB = L
    return _LIB


def _P(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def compact_line(kind: int, h: float, f: np.ndarray) -> np.ndarray:
    """Compact first derivative of one line: f holds nodes 0..m+1; returns
    rows 1..m (kind 0: equally spaced ends, kind 1: walls half a spacing
    away)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    m = f.size - 2
    out = np.zeros(m + 2)
    lib().vt_compact_line(kind, m, h, _P(f), _P(out))
    return out[1:m + 1]


class Tank:
    """One state of the tank: its parameters and 21 fields."""

    def __init__(self, nr, nz, np_, dt, Re, Ptl, Gr, dr=None, dz=None, dp=None, lx=1.0, lz=1.0, ly=4.0):
        dr = dr if dr is not None else lx / nr
        dz = dz if dz is not None else lz / nz
        dp = dp if dp is not None else ly / np_
        self.p = VtParams(nr, nz, np_, nr + 2, nz + 2, dt, dr, dz, dp, Re, Ptl, Gr)
        self.shape = (np_ + 2, nz + 2, nr + 2)
        self.f = {n: np.zeros(self.shape) for n in FIELDS}
        self.c = VtFields(*[_P(self.f[n]) for n in FIELDS])
        self.d = [np.zeros(self.shape) for _ in range(12)]
        self.dc = VtDerivs((_D * 12)(*[_P(a) for a in self.d]))
        self._lu = {}

    def __getitem__(self, name):
        return self.f[name]

    def __setitem__(self, name, val):
        self.f[name][...] = val

    def _b(self):
        return ctypes.byref(self.p), ctypes.byref(self.c)

    def derivatives(self):
        lib().vt_derivatives(ctypes.byref(self.p), ctypes.byref(self.c), ctypes.byref(self.dc))
        return self.d

    def call(self, name):
        if name in ("avorticity", "bvorticity", "cvorticity", "temperature"):
            getattr(lib(), "vt_" + name)(*self._b(), ctypes.byref(self.dc))
        else:
            getattr(lib(), "vt_" + name)(*self._b())

    def heat(self):
        out = np.zeros(2)
        lib().vt_heat(*self._b(), _P(out))
        return float(out[0]), float(out[1])

    def energy(self):
        return lib().vt_energy(*self._b())

    def ranges(self, which):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib().vt_ranges(ctypes.byref(self.p), which, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def apply(self, which, psi):
        out = np.zeros(self.shape)
        lib().vt_apply(ctypes.byref(self.p), which, _P(np.ascontiguousarray(psi)), _P(out))
        return out

    def residual(self, which, psi, rhs):
        res = np.zeros(self.shape)
        m = lib().vt_residual(ctypes.byref(self.p), which, _P(np.ascontiguousarray(psi)),
                              _P(np.ascontiguousarray(rhs)), _P(res))
        return m, res

    def psi_bounds(self, which, psi):
        lib().vt_psi_bounds(ctypes.byref(self.p), which, _P(psi))

    def matrix(self, which):
        """The operator of `which` as a dense matrix over its unknowns
        (column q = L applied to unit vector q), with the index arrays."""
        i1, j1, k1 = self.ranges(which)
        idx = [(k, j, i) for k in range(1, k1 + 1) for j in range(1, j1 + 1) for i in range(1, i1 + 1)]
        n = len(idx)
        M = np.zeros((n, n))
        e = np.zeros(self.shape)
        kk = np.array([t[0] for t in idx]), np.array([t[1] for t in idx]), np.array([t[2] for t in idx])
        for q, (k, j, i) in enumerate(idx):
            e[k, j, i] = 1.0
            M[:, q] = self.apply(which, e)[kk]
            e[k, j, i] = 0.0
        return M, kk

    def solve(self, which, rhs):
        """Exact solution of L psi = rhs (dense LU, small grids only), with
        the output conventions of a solve (walls 0, y ghosts mirrored)."""
        import scipy.linalg as sl
        if which not in self._lu:
            M, kk = self.matrix(which)
            self._lu[which] = (sl.lu_factor(M), kk)
        lu, kk = self._lu[which]
        psi = np.zeros(self.shape)
        psi[kk] = sl.lu_solve(lu, rhs[kk])
        self.psi_bounds(which, psi)
        return psi

    def poisson(self, which):
        """poisson1/2/3 with the reference's conventions: psip <- psi, psi
        <- the exact solution (whole field: walls and ghosts as a solve
        writes them)."""
        psi, psip, rhs = PSI[which]
        self.f[psip][...] = self.f[psi]
        self.f[psi][...] = self.solve(which, self.f[rhs])

    def step(self):
        """One time step in the order of rot3dmg.f:241-253; returns the
        heat fluxes."""
        self.derivatives()
        for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
            self.call(rt)
        h = self.heat()
        for which, rhs in ((1, "rhsw"), (2, "rhsv"), (3, "rhsu")):
            self.call(rhs)
            self.poisson(which)
        self.call("update")
        return h
