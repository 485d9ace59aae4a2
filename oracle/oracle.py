"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle is a plain-C restatement of velvor2dperiodicrbconv.f (vv_oracle.c).
Only tests/, the smoke entry point and bench.py's cpu_baseline leg use it, as
the checker.  Fields are numpy arrays of shape (np+1, ld) in C order, so that
F[k, i] is the reference's F(i,1,k) with leading dimension ld.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.vo_residual.restype = ctypes.c_double
        L.vo_residual_g.restype = ctypes.c_double
        L.vo_energy.restype = ctypes.c_double
        _lib = L
    return _lib


def P(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


class VoParams(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("np", ctypes.c_int), ("ld", ctypes.c_int),
                ("dt", ctypes.c_double), ("dr", ctypes.c_double), ("dp", ctypes.c_double),
                ("Ra", ctypes.c_double), ("Pr", ctypes.c_double), ("geom", ctypes.c_int),
                ("r0", ctypes.c_double), ("Re", ctypes.c_double), ("robert", ctypes.c_double)]


_FIELD_NAMES = ["C", "Cp", "Cn", "th", "thp", "thn", "uel", "vel", "wel",
                "uelp", "velp", "A", "B", "An", "Bn", "Ap"]


class VoFields(ctypes.Structure):
    _fields_ = [(n, _D) for n in _FIELD_NAMES]


MAXLEV = 8


class VoMG(ctypes.Structure):
    _fields_ = [("nlev", ctypes.c_int), ("ld", ctypes.c_int), ("ncols", ctypes.c_int),
                ("n1", ctypes.c_int * MAXLEV), ("n3", ctypes.c_int * MAXLEV),
                ("hr", ctypes.c_double * MAXLEV), ("hp", ctypes.c_double * MAXLEV),
                ("hhr", ctypes.c_double * MAXLEV), ("hhp", ctypes.c_double * MAXLEV),
                ("v", _D * MAXLEV), ("f", _D * MAXLEV),
                ("psif", _D), ("psic", _D), ("rc", _D), ("rf", _D), ("tempa", _D),
                ("nlev_cap", ctypes.c_int), ("geom", ctypes.c_int), ("r0", ctypes.c_double),
                ("nsweep", ctypes.c_int), ("zero_seam", ctypes.c_int)]


class Run:
    """One run held by the oracle (mirror of the reference's COMMON state,
    rb.f:69-76).  geom 0: periodic box of velvor2dperiodicrbconv.f (width
    given, dp = width/np); geom 1: rotating-annulus plane of
    velvor2dnewsch3.4.f (dp = 2 pi/np, ann.f:124), with Re, Robert 0.05,
    five multigrid levels at most and three sweeps per relaxation."""

    def __init__(self, nr, np_, dt, Ra, Pr, width=3.0, height=0.5, ld=None, nlev_cap=None,
                 geom=0, r0=0.5, Re=0.0, robert=None):
        self.nr, self.np = nr, np_
        self.geom = geom
        self.ld = ld if ld is not None else nr + 2
        dr = height / nr
        dp = width / np_ if geom == 0 else 2.0 * np.arccos(-1.0) / np_
        if robert is None:
            robert = 0.02 if geom == 0 else 0.05
        if nlev_cap is None:
            nlev_cap = 4 if geom == 0 else 5
        self.p = VoParams(nr, np_, self.ld, dt, dr, dp, Ra, Pr, geom, r0, Re, robert)
        self.arr = {n: np.zeros((np_ + 1, self.ld)) for n in _FIELD_NAMES}
        self.vrh = np.zeros((np_ + 1, self.ld))
        self.urh = np.zeros((np_ + 1, self.ld))
        self.f = VoFields(*[P(self.arr[n]) for n in _FIELD_NAMES])
        self.mg = VoMG()
        lib().vo_mg_init(ctypes.byref(self.mg), nr, np_, ctypes.c_double(dr),
                         ctypes.c_double(dp), nlev_cap)
        if geom == 1:
            lib().vo_mg_set_geom(ctypes.byref(self.mg), 1, ctypes.c_double(r0), 3, 0)
        self.kt = 0

    def __del__(self):
        try:
            lib().vo_mg_free(ctypes.byref(self.mg))
        except Exception:
            pass

    def __getitem__(self, name):
        return self.arr[name]

    def seed(self, r0=0.5, amp=1e-3, seed=20261202):
        lib().vo_seed_temperature(ctypes.byref(self.p), P(self.arr["th"]),
                                  ctypes.c_double(r0), ctypes.c_double(amp),
                                  ctypes.c_ulonglong(seed))
        self.start()

    def start(self):
        """rb.f:315-329: previous levels copy the current ones."""
        for a, b in (("Cp", "C"), ("thp", "th"), ("uelp", "uel"), ("velp", "vel")):
            self.arr[a][...] = self.arr[b]
        self.kt = 0

    def step(self, mode=1, tolcon=None):
        """mode 0: literal reference Poisson (absolute stop tolcon: 2e-6 in
        rb.f:2086, 1e-5 in ann.f:2093); mode 1: exact Poisson."""
        if tolcon is None:
            tolcon = 2e-6 if self.geom == 0 else 10e-6
        trans = np.zeros(2)
        ncl = ctypes.c_int(0)
        first = 1 if self.kt == 0 else 0
        st = lib().vo_step(ctypes.byref(self.p), ctypes.byref(self.f), ctypes.byref(self.mg),
                           P(self.vrh), P(self.urh), first, mode, ctypes.c_double(tolcon),
                           P(trans), ctypes.byref(ncl))
        self.kt += 1
        return st, trans, ncl.value

    def nusselt(self, trans, width):
        height = self.p.dr * self.nr
        return trans * height / width

    def energy(self):
        """Kinetic energy of uel, vel (the build's definition, vv_oracle.h)."""
        return lib().vo_energy(ctypes.byref(self.p), P(self.arr["uel"]), P(self.arr["vel"]))


def tdma(a, b, c, r):
    a, b, c, r = (np.array(x, dtype=np.float64) for x in (a, b, c, r))
    y = np.zeros_like(r)
    lib().vo_tdma(P(a), P(b), P(c), P(y), P(r), len(r))
    return y


def ptdma(a, b, c, r):
    a, b, c, r = (np.array(x, dtype=np.float64) for x in (a, b, c, r))
    y = np.zeros_like(r)
    lib().vo_ptdma(P(a), P(b), P(c), P(y), P(r), len(r))
    return y


def pdma(a, b, c, d, e, r):
    a, b, c, d, e, r = (np.array(x, dtype=np.float64) for x in (a, b, c, d, e, r))
    x = np.zeros_like(r)
    lib().vo_pdma(P(a), P(b), P(c), P(d), P(e), P(r), P(x), len(r))
    return x


def poisson_exact(which, psi, psip, rhs, nr, np_, dr, dp, geom=0, r0=0.0):
    ld = psi.shape[1]
    lib().vo_poisson_exact_g(which, P(psi), P(psip), P(rhs), nr, np_, ld,
                             ctypes.c_double(dr), ctypes.c_double(dp), geom, ctypes.c_double(r0))


def residual(which, psi, rhs, nr, np_, dr, dp, out=None, geom=0, r0=0.0):
    ld = psi.shape[1]
    o = out if out is not None else np.zeros_like(psi)
    m = lib().vo_residual_g(which, P(psi), P(rhs), P(o), nr, np_, ld,
                            ctypes.c_double(dr), ctypes.c_double(dp), geom, ctypes.c_double(r0))
    return m, o


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    """U[0,1) doubles from splitmix64 (same stream as vo_seed_temperature)."""
    out = np.empty(n)
    s = seed & 0xFFFFFFFFFFFFFFFF
    M = 0xFFFFFFFFFFFFFFFF
    for j in range(n):
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z = z ^ (z >> 31)
        out[j] = (z >> 11) * (1.0 / 9007199254740992.0)
    return out
