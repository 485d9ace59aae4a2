"""In-process driver for the reference build oracle/_ref/librb.so.

librb.so is velvor2dperiodicrbconv.f compiled unmodified by oracle/Makefile
from the reference sources (only in the container that holds them).  Its
subroutines take every argument by reference and keep their state in COMMON
blocks, which are exported as the symbols sol1_, sol2_, sol3_, num_, val_,
radius_.  This module maps those blocks onto numpy arrays so a test can set
a state, call one reference subroutine and read the result, and it replays
the reference time loop (rb.f:341-382) step by step.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBRB = os.path.join(ROOT, "oracle", "_ref", "librb.so")
LIBANN = os.path.join(ROOT, "oracle", "_ref", "libann.so")

KR1, KP1 = 291, 1611          # rb.f:54 parameter(kr=290, kp=1610)
ANN_KR1, ANN_KP1 = 421, 3601  # velvor2dnewsch3.4.f: parameter(kr=420, kp=3600)
_D = ctypes.POINTER(ctypes.c_double)


def available(kind="rb") -> bool:
    return os.path.exists(LIBRB if kind == "rb" else LIBANN)


class Reference:
    """Singleton view of one reference build's COMMON state
    (kind "rb": velvor2dperiodicrbconv.f, "ann": velvor2dnewsch3.4.f)."""

    _inst = {}

    def __new__(cls, kind="rb"):
        if kind not in cls._inst:
            inst = super().__new__(cls)
            inst._init(kind)
            cls._inst[kind] = inst
        return cls._inst[kind]

    def _init(self, kind):
        self.kind = kind
        L = ctypes.CDLL(LIBRB if kind == "rb" else LIBANN)
        self.L = L
        self.KR1, self.KP1 = (KR1, KP1) if kind == "rb" else (ANN_KR1, ANN_KP1)
        KR1_, KP1_ = self.KR1, self.KP1
        nf = KR1_ * KP1_

        def block(name, n):
            addr = ctypes.addressof(ctypes.c_char.in_dll(L, name))
            buf = (ctypes.c_double * (n * nf)).from_address(addr)
            return np.frombuffer(buf, dtype=np.float64).reshape(n, KP1_, KR1_)

        s1 = block("sol1_", 9)
        s2 = block("sol2_", 3)
        s3 = block("sol3_", 6)
        self.f = {}
        for j, n in enumerate(["An", "A", "Ap", "Bn", "B", "Bp", "Cn", "C", "Cp"]):
            self.f[n] = s1[j]
        for j, n in enumerate(["thn", "th", "thp"]):
            self.f[n] = s2[j]
        for j, n in enumerate(["uel", "vel", "wel", "uelp", "velp", "welp"]):
            self.f[n] = s3[j]
        numaddr = ctypes.addressof(ctypes.c_char.in_dll(L, "num_"))
        self.num_d = (ctypes.c_double * 4).from_address(numaddr)
        self.num_i = (ctypes.c_int * 3).from_address(numaddr + 32)
        self.val = (ctypes.c_double * 4).in_dll(L, "val_")
        self.radius = (ctypes.c_double * 2).in_dll(L, "radius_")
        self.vrh = np.zeros((KP1_, KR1_))
        self.urh = np.zeros((KP1_, KR1_))

    # ------------------------------------------------------------ setup
    def setup(self, nr, np_, dt, Ra, Pr, r0=0.5, rmax=1.0, width=3.0, Re=0.0):
        """Scalars as the main program sets them (rb.f:78-133; the annulus
        program uses dp = 2 pi/np, ann.f:124)."""
        dr = (rmax - r0) / nr
        dp = width / np_ if self.kind == "rb" else 2.0 * np.arccos(-1.0) / np_
        self.num_d[0], self.num_d[1], self.num_d[2], self.num_d[3] = dt, dr, 0.0, dp
        self.num_i[0], self.num_i[1], self.num_i[2] = nr, 1, np_
        self.val[0], self.val[1], self.val[2], self.val[3] = Ra, Re, Pr, 0.0
        self.radius[0], self.radius[1] = r0, rmax
        for a in self.f.values():
            a[...] = 0.0
        self.nr, self.np = nr, np_
        self.dr, self.dp = dr, dp
        n = ctypes.c_int
        d = ctypes.c_double
        self.nlev = n(0)
        self.L.poissoninit_(ctypes.byref(n(nr)), ctypes.byref(n(1)), ctypes.byref(n(np_)),
                            ctypes.byref(d(dr)), ctypes.byref(d(0.0)), ctypes.byref(d(dp)),
                            ctypes.byref(self.nlev))
        self.kt = 0

    def load(self, name, arr):
        """arr: (np+1, ni) -> COMMON field rows i=0..ni-1."""
        a = self.f[name]
        a[: arr.shape[0], : arr.shape[1]] = arr

    def get(self, name, ni=None):
        ni = ni if ni is not None else self.nr + 2
        return np.array(self.f[name][: self.np + 1, :ni])

    def start(self):
        """rb.f:315-329"""
        for a, b in (("Ap", "A"), ("Bp", "B"), ("Cp", "C"), ("thp", "th"),
                     ("uelp", "uel"), ("velp", "vel")):
            self.f[a][...] = self.f[b]
        self.f["welp"][...] = 0.0
        self.kt = 0

    # ------------------------------------------------------ subroutines
    def call(self, name, *args):
        getattr(self.L, name + "_")(*args)

    def poisson(self, which, psi, psip, rhs):
        """Call poisson2/poisson3 on arrays of shape (KP1, KR1) of this build."""
        n = ctypes.c_int
        d = ctypes.c_double
        P = lambda a: a.ctypes.data_as(_D)
        fn = self.L.poisson2_ if which == 2 else self.L.poisson3_
        fn(P(psi), P(psip), P(rhs), ctypes.byref(n(self.nr)), ctypes.byref(n(1)),
           ctypes.byref(n(self.np)), ctypes.byref(d(self.dr)), ctypes.byref(d(0.0)),
           ctypes.byref(d(self.dp)), ctypes.byref(self.nlev))

    def heat(self):
        t1, t2 = ctypes.c_double(0.0), ctypes.c_double(0.0)
        self.L.heat_(ctypes.byref(t1), ctypes.byref(t2))
        return t1.value, t2.value

    def step(self):
        """One pass of the time loop body in the reference's order."""
        first = self.kt == 0
        self.call("cvorticity1" if first else "cvorticity")
        self.call("temperature1" if first else "temperature")
        t = self.heat()
        P = lambda a: a.ctypes.data_as(_D)
        self.L.rhsv_(P(self.vrh))
        self.poisson(2, self.f["vel"], self.f["velp"], self.vrh)
        self.L.rhsu_(P(self.urh))
        self.poisson(3, self.f["uel"], self.f["uelp"], self.urh)
        self.call("update")
        self.kt += 1
        return t
