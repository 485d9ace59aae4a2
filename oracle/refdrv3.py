"""In-process driver for the 3D reference builds oracle/_ref/librot3d.so and
librot3dhor.so (test and CPU-baseline infrastructure; no product code
imports it).

librot3d.so is rot3dmg.f, librot3dhor.so rot3dmghor.f (the stationary
horizontal annulus), each compiled unmodified by oracle/Makefile from the
reference sources (only in the container that holds them); both declare the
same COMMON blocks.  Its COMMON
blocks (/sol1/ /sol2/ /sol3/ /num/ /val/ /radius/, and the multigrid
/arrays/ /backup/ /step/ /ints/) are exported symbols; this module maps them
onto numpy arrays with the reference's PARAMETER extents (kr=kz=33,
kp=160, mgen=202590; rot3dmg.f:23) so a test can set a state, call one
reference subroutine and read the result.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_ref", "librot3d.so")
LIBS = {0: LIB, 1: os.path.join(ROOT, "oracle", "_ref", "librot3dhor.so")}
KR1, KZ1, KP1 = 34, 34, 161
MGEN = 202590
SHAPE = (KP1, KZ1, KR1)
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def available(model: int = 0) -> bool:
    return os.path.exists(LIBS[model])


class Reference3:
    """The reference build of model 0 (rot3dmg.f) or 1 (rot3dmghor.f); one
    instance per library, since its COMMON blocks are process-wide."""
    _inst = {}

    def __new__(cls, model: int = 0):
        if model not in cls._inst:
            inst = super().__new__(cls)
            inst.model = model
            inst._init()
            cls._inst[model] = inst
        return cls._inst[model]

    def _init(self):
        L = ctypes.CDLL(LIBS[self.model])
        self.L = L
        nf = KR1 * KZ1 * KP1

        def block(name, n, count=None):
            addr = ctypes.addressof(ctypes.c_char.in_dll(L, name))
            buf = (ctypes.c_double * (n * nf)).from_address(addr)
            return np.frombuffer(buf, dtype=np.float64).reshape((n,) + SHAPE)

        s1, s2, s3, ar = block("sol1_", 9), block("sol2_", 3), block("sol3_", 6), block("arrays_", 5)
        self.f = {}
        for j, n in enumerate(["An", "A", "Ap", "Bn", "B", "Bp", "Cn", "C", "Cp"]):
            self.f[n] = s1[j]
        for j, n in enumerate(["thn", "th", "thp"]):
            self.f[n] = s2[j]
        for j, n in enumerate(["uel", "vel", "wel", "uelp", "velp", "welp"]):
            self.f[n] = s3[j]
        for j, n in enumerate(["psic", "psif", "rc", "rf", "tempa"]):
            self.f[n] = ar[j]
        baddr = ctypes.addressof(ctypes.c_char.in_dll(L, "backup_"))
        bk = np.frombuffer((ctypes.c_double * (2 * MGEN)).from_address(baddr), dtype=np.float64)
        self.v, self.fl = bk[:MGEN], bk[MGEN:]
        numaddr = ctypes.addressof(ctypes.c_char.in_dll(L, "num_"))
        self.num_d = (ctypes.c_double * 4).from_address(numaddr)
        self.num_i = (ctypes.c_int * 3).from_address(numaddr + 32)
        self.val = (ctypes.c_double * 4).in_dll(L, "val_")
        self.radius = (ctypes.c_double * 2).in_dll(L, "radius_")
        self.ints = (ctypes.c_int * 24).in_dll(L, "ints_")
        for n in ("urh", "vrh", "wrh"):
            self.f[n] = np.zeros(SHAPE)

    def setup(self, p):
        """Scalars from a V3Params (the main program's rot3dmg.f:73-93)."""
        self.num_d[:] = [p.dt, p.dr, p.dz, p.dp]
        self.num_i[:] = [p.nr, p.nz, p.np]
        self.val[:] = [p.Ra, p.Re, p.Ptl, p.Gr]
        self.radius[:] = [p.r0, p.rmax]

    def call(self, name):
        getattr(self.L, name + "_")()

    def heat(self):
        return self.heat2()[0]

    def heat2(self):
        """(trans1, trans2): heat(trans1) of rot3dmg.f (trans2 = 0), or
        heat(trans1,trans2) of rot3dmghor.f (rot3dmghor.f:913)."""
        t1, t2 = ctypes.c_double(0.0), ctypes.c_double(0.0)
        if self.model == 1:
            self.L.heat_(ctypes.byref(t1), ctypes.byref(t2))
        else:
            self.L.heat_(ctypes.byref(t1))
        return t1.value, t2.value

    def rhs(self, which):
        name = {1: "wrh", 2: "vrh", 3: "urh"}[which]
        fn = {1: "rhsw_", 2: "rhsv_", 3: "rhsu_"}[which]
        getattr(self.L, fn)(self.f[name].ctypes.data_as(_D))

    def relax_residual(self, which, p, psi, rhs):
        """One pass of relax1/2/3 on the finest level (rot3dmg.f:1543,
        1739, 1986): psi and rhs go through trans1 into the flat
        multigrid storage, the routine relaxes once and stores the residual
        of the relaxed iterate in /arrays/ rc.  Returns (psif, rc)."""
        nlev = ctypes.c_int(0)
        ni, nz, np_ = ctypes.c_int(p.nr), ctypes.c_int(p.nz), ctypes.c_int(p.np)
        dr, dz, dp = ctypes.c_double(p.dr), ctypes.c_double(p.dz), ctypes.c_double(p.dp)
        self.L.poissoninit_(ctypes.byref(ni), ctypes.byref(nz), ctypes.byref(np_), ctypes.byref(dr),
                            ctypes.byref(dz), ctypes.byref(dp), ctypes.byref(nlev))
        ibeg, n1, n2, n3 = self.ints[0], self.ints[6], self.ints[12], self.ints[18]
        self.v[:] = 0.0
        self.fl[:] = 0.0
        self.f["rf"][...] = rhs
        self.f["psif"][...] = psi
        args = [ctypes.byref(ctypes.c_int(x)) for x in (ibeg, n1, n2, n3)]
        self.L.trans1_(self.f["rf"].ctypes.data_as(_D), self.fl.ctypes.data_as(_D), *args)
        self.L.trans1_(self.f["psif"].ctypes.data_as(_D), self.v.ctypes.data_as(_D), *args)
        nrel, klev, nconv = ctypes.c_int(1), ctypes.c_int(1), ctypes.c_int(0)
        fn = {1: "relax1_", 2: "relax2_", 3: "relax3_"}[which]
        getattr(self.L, fn)(ctypes.byref(nrel), ctypes.byref(klev), ctypes.byref(nconv), ctypes.byref(nlev))
        return self.f["psif"].copy(), self.f["rc"].copy()

    def poissoninit(self, p) -> int:
        """poissoninit (rot3dmg.f:1153): the multigrid levels; returns nlev."""
        nlev = ctypes.c_int(0)
        ni, nz, np_ = ctypes.c_int(p.nr), ctypes.c_int(p.nz), ctypes.c_int(p.np)
        dr, dz, dp = ctypes.c_double(p.dr), ctypes.c_double(p.dz), ctypes.c_double(p.dp)
        self.L.poissoninit_(ctypes.byref(ni), ctypes.byref(nz), ctypes.byref(np_), ctypes.byref(dr),
                            ctypes.byref(dz), ctypes.byref(dp), ctypes.byref(nlev))
        return nlev.value

    def step(self, p, nlev) -> float:
        """One pass of the reference time loop (rot3dmg.f:241-253) with the
        reference's own multigrid solves; returns trans1."""
        for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
            self.call(rt)
        t = self.heat()
        refs = lambda: [ctypes.byref(ctypes.c_int(p.nr)), ctypes.byref(ctypes.c_int(p.nz)),
                        ctypes.byref(ctypes.c_int(p.np)), ctypes.byref(ctypes.c_double(p.dr)),
                        ctypes.byref(ctypes.c_double(p.dz)), ctypes.byref(ctypes.c_double(p.dp)),
                        ctypes.byref(ctypes.c_int(nlev))]
        for which, (psi, psip, rhs) in ((1, ("wel", "welp", "wrh")), (2, ("vel", "velp", "vrh")),
                                        (3, ("uel", "uelp", "urh"))):
            self.rhs(which)
            getattr(self.L, f"poisson{which}_")(self.f[psi].ctypes.data_as(_D), self.f[psip].ctypes.data_as(_D),
                                                 self.f[rhs].ctypes.data_as(_D), *refs())
        self.call("update")
        return t


def time_reference_steps(state_npz: str, out_json: str, nsteps: int, budget_s: float) -> None:
    """Times the reference build's time step on the state saved in
    state_npz (scalars nr, nz, np, dt, Ra, Re, Ptl, Gr, r0, rmax, zmax and
    the fields) and writes {"steps", "seconds", "trans1"} to out_json.  The
    reference prints to stdout, so this runs in a process of its own."""
    import json
    import oracle3 as O3
    z = np.load(state_npz)
    nr, nz, np_ = int(z["nr"]), int(z["nz"]), int(z["np"])
    run = O3.Run3(nr, nz, np_, float(z["dt"]), float(z["Ra"]), float(z["Re"]), float(z["Ptl"]), float(z["Gr"]),
                  r0=float(z["r0"]), rmax=float(z["rmax"]), zmax=float(z["zmax"]))
    ref = Reference3()
    ref.setup(run.p)
    K = np_ + 1
    for n in O3.FIELDS:
        ref.f[n][...] = 0.0
        if "f_" + n in z:
            ref.f[n][:K, :nz + 2, :nr + 2] = z["f_" + n]
    nlev = ref.poissoninit(run.p)
    import time
    n, t = 0, 0.0
    t0 = time.perf_counter()
    while n < nsteps:
        t = ref.step(run.p, nlev)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    with open(out_json, "w") as f:
        json.dump({"steps": n, "seconds": el, "trans1": t, "nlev": nlev}, f)


if __name__ == "__main__":
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    time_reference_steps(sys.argv[1], sys.argv[2], int(sys.argv[3]), float(sys.argv[4]))
