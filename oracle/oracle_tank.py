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


# ---------------------------------------------------------------- fast solve
# The exact solve of the same operators at sizes the dense LU cannot reach:
# along y the cosine modes (u, w) or sine modes (v) by scipy's DCT-II / DST-I,
# along z and x the eigenbases of the operator's rows (numpy), then a
# division by the sum of the three eigenvalues.  Pinned against the dense
# solve in tests/test_oracle_tank.py.

def _rows_matrix(kind: int, n: int, h2: float) -> np.ndarray:
    """The rows of one direction as vt_apply forms them (kind 0 Dirichlet,
    1 the no-slip closure, 2 mirrored ghosts)."""
    M = np.zeros((n, n))
    for t in range(n):
        M[t, t] = -2.0
        if t > 0:
            M[t, t - 1] = 1.0
        if t < n - 1:
            M[t, t + 1] = 1.0
    if kind == 1:
        M[0, :3] = [-2.0, 5.0 / 3.0, -0.2]
        M[-1, -3:] = [-0.2, 5.0 / 3.0, -2.0]
    elif kind == 2:
        M[0, 0] = M[-1, -1] = -1.0
    return M / h2


def _kinds(which):
    return (0 if which == 3 else 1), (0 if which == 1 else 1), (0 if which == 2 else 2)


class FastSolver:
    """Exact separable solves of a Tank's three operators."""

    def __init__(self, tank: "Tank"):
        self.t = tank
        self.ops = {}
        p = tank.p
        for which in (1, 2, 3):
            i1, j1, k1 = tank.ranges(which)
            kx, kz, ky = _kinds(which)
            lx, Vx = np.linalg.eig(_rows_matrix(kx, i1, p.dr * p.dr))
            lz, Vz = np.linalg.eig(_rows_matrix(kz, j1, p.dz * p.dz))
            assert np.max(np.abs(lx.imag)) == 0.0 and np.max(np.abs(lz.imag)) == 0.0
            lx, Vx, lz, Vz = lx.real, Vx.real, lz.real, Vz.real
            q = np.arange(1, k1 + 1) if which == 2 else np.arange(k1)
            ly = (2.0 * np.cos(np.pi * q / p.np) - 2.0) / (p.dp * p.dp)
            self.ops[which] = (lx, Vx, np.linalg.inv(Vx), lz, Vz, np.linalg.inv(Vz), ly)

    def solve(self, which, rhs):
        import scipy.fft as sf
        t = self.t
        i1, j1, k1 = t.ranges(which)
        lx, Vx, Vxi, lz, Vz, Vzi, ly = self.ops[which]
        f = rhs[1:k1 + 1, 1:j1 + 1, 1:i1 + 1]
        s = sf.dst(f, type=1, axis=0) if which == 2 else sf.dct(f, type=2, axis=0)
        s = np.matmul(Vzi, s)                       # along j
        s = np.matmul(s, Vxi.T)                     # along i
        s /= (ly[:, None, None] + lz[None, :, None] + lx[None, None, :])
        s = np.matmul(s, Vx.T)
        s = np.matmul(Vz, s)
        x = sf.idst(s, type=1, axis=0) if which == 2 else sf.idct(s, type=2, axis=0)
        psi = np.zeros(t.shape)
        psi[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = x
        t.psi_bounds(which, psi)
        return psi


def time_step(npz_in: str, json_out: str) -> None:
    """A bounded sample of one step of the oracle on the state in npz_in, on
    one thread: the explicit routines of the whole step in C (derivatives,
    A/B/C/T updates, heat, the three right-hand sides, update) and one of its
    three exact fast solves (all three cost the same: the transforms of
    equal-sized operators).  Writes the seconds of each part and the step
    time they give (explicit + 3 solves)."""
    import json
    import time
    from threadpoolctl import threadpool_limits
    z = np.load(npz_in)
    nr, nz, np_ = int(z["nr"]), int(z["nz"]), int(z["np"])
    t = Tank(nr, nz, np_, float(z["dt"]), float(z["Re"]), float(z["Ptl"]), float(z["Gr"]), lx=float(z["lx"]),
             lz=float(z["lz"]), ly=float(z["ly"]))
    for n in FIELDS:
        if "f_" + n in z:
            t[n] = z["f_" + n]
    with threadpool_limits(limits=1):
        fs = FastSolver(t)
        t0 = time.perf_counter()
        t.derivatives()
        for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
            t.call(rt)
        h = t.heat()
        for rhs in ("rhsw", "rhsv", "rhsu"):
            t.call(rhs)
        t.call("update")
        t1 = time.perf_counter()
        psi, psip, f = PSI[2]
        t[psip] = t[psi]
        t[psi] = fs.solve(2, t[f])
        t2 = time.perf_counter()
    with open(json_out, "w") as fo:
        json.dump({"explicit_s": t1 - t0, "solve_s": t2 - t1, "step_s": (t1 - t0) + 3.0 * (t2 - t1),
                   "heat": list(h)}, fo)


if __name__ == "__main__":
    import sys
    time_step(sys.argv[1], sys.argv[2])
