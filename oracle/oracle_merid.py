"""ctypes wrapper of the meridional-plane oracle (vvm_oracle.c in liboracle.so)
and its exact Poisson solves in numpy.

Test infrastructure only: the tests, the smoke check and the benchmark's CPU
leg use it; no product code imports it.  Fields are numpy arrays of shape
(nz+1, nr+1) in C order: X[j, i] is X(r_i, z_j).

The physics is the velocity-vorticity form of rotann2.f's meridional plane
(azimuthal vorticity, swirl, temperature; rotann2.f:268-420) on tanh-clustered
nodes, written out in vvm_oracle.h.
"""
from __future__ import annotations

import ctypes
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_D = ctypes.POINTER(ctypes.c_double)

FIELDS = ["w", "wp", "wn", "v", "vp", "vn", "th", "thp", "thn", "u", "ww", "fu", "fw"]
DNAMES = ["w_r", "w_z", "v_r", "v_z", "th_r", "th_z"]


class VmParams(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nz", ctypes.c_int), ("dt", ctypes.c_double), ("Ra", ctypes.c_double),
                ("Re", ctypes.c_double), ("Pr", ctypes.c_double), ("r", _D), ("z", _D)]


class VmFields(ctypes.Structure):
    _fields_ = [(n, _D) for n in FIELDS]


def lib():
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(os.path.join(HERE, "liboracle.so"))
        for fn in ("vm_fornberg", "vm_nodes", "vm_compact_line", "vm_derivatives", "vm_explicit", "vm_rhs",
                   "vm_apply", "vm_update", "vm_heat"):
            getattr(L, fn).restype = None
        L.vm_fornberg.argtypes = [_D, ctypes.c_int, ctypes.c_double, ctypes.c_int, _D]
        L.vm_nodes.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D]
        L.vm_compact_line.argtypes = [ctypes.c_int, _D, _D, _D]
        _LIB = L
    return _LIB


def _P(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def fornberg(x, xx: float, m: int) -> np.ndarray:
    """Weights c[k, q] of node q for the k-th derivative at xx (wfdif.f:59-85)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    c = np.zeros((m + 1, x.size))
    lib().vm_fornberg(_P(x), x.size, float(xx), m, _P(c))
    return c


def nodes(n: int, a: float, b: float, beta: float) -> np.ndarray:
    """n+1 tanh-clustered nodes on [a, b] (beta = 0: uniform)."""
    x = np.zeros(n + 1)
    lib().vm_nodes(n, a, b, beta, _P(x))
    return x


def compact_line(x, f) -> np.ndarray:
    """Compact first derivative of f on the nodes x at the inner nodes
    1..n-1."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros_like(f)
    lib().vm_compact_line(x.size - 1, _P(x), _P(f), _P(out))
    return out[1:-1]


def three_point(x, i: int):
    """Fornberg's three-point weights at inner node i: (first, second)."""
    c = fornberg(x[i - 1:i + 2], x[i], 2)
    return c[1], c[2]


class Merid:
    """One state of the meridional plane: parameters, nodes and fields.

    r0, rmax: the cylinders; H: the height; beta: the tanh clustering
    toward all four walls."""

    def __init__(self, nr, nz, dt, Ra, Re, Pr, r0=0.5, rmax=1.0, H=0.5, beta=1.5, r=None, z=None):
        self.r = nodes(nr, r0, rmax, beta) if r is None else np.ascontiguousarray(r, dtype=np.float64)
        self.z = nodes(nz, 0.0, H, beta) if z is None else np.ascontiguousarray(z, dtype=np.float64)
        self.p = VmParams(nr, nz, dt, Ra, Re, Pr, _P(self.r), _P(self.z))
        self.nr, self.nz = nr, nz
        self.shape = (nz + 1, nr + 1)
        self.f = {n: np.zeros(self.shape) for n in FIELDS}
        self.c = VmFields(*[_P(self.f[n]) for n in FIELDS])
        self.d = [np.zeros(self.shape) for _ in range(6)]
        self._dp = (_D * 6)(*[_P(a) for a in self.d])
        self._solver = None

    def __getitem__(self, k):
        return self.f[k]

    def set(self, **kw):
        for k, v in kw.items():
            self.f[k][...] = v

    # the routines of the step
    def derivatives(self):
        lib().vm_derivatives(ctypes.byref(self.p), ctypes.byref(self.c), self._dp)

    def explicit(self):
        lib().vm_explicit(ctypes.byref(self.p), ctypes.byref(self.c), self._dp)

    def rhs(self):
        lib().vm_rhs(ctypes.byref(self.p), ctypes.byref(self.c))

    def apply(self, which: int, psi) -> np.ndarray:
        psi = np.ascontiguousarray(psi, dtype=np.float64)
        out = np.zeros(self.shape)
        lib().vm_apply(ctypes.byref(self.p), which, _P(psi), _P(out))
        return out

    def update(self):
        lib().vm_update(ctypes.byref(self.p), ctypes.byref(self.c))

    def heat(self):
        out = np.zeros(2)
        lib().vm_heat(ctypes.byref(self.p), ctypes.byref(self.c), _P(out))
        return out

    # the Poisson operators and their exact solves
    def rows(self, axis: str, which: int = 2):
        """Tridiagonal rows (sub, diag, sup) of the r part (axis 'r', with
        -1/r^2 for which = 1) or the z part of the operator on the inner
        nodes."""
        x = self.r if axis == "r" else self.z
        n = x.size - 1
        sub, dia, sup = np.zeros(n - 1), np.zeros(n - 1), np.zeros(n - 1)
        for i in range(1, n):
            c1, c2 = three_point(x, i)
            if axis == "r":
                rr = x[i]
                w = c2 + c1 / rr
                sub[i - 1], dia[i - 1], sup[i - 1] = w[0], w[1], w[2]
                if which == 1:
                    dia[i - 1] -= 1.0 / (rr * rr)
            else:
                sub[i - 1], dia[i - 1], sup[i - 1] = c2
        return sub, dia, sup

    def matrix(self, which: int):
        """The operator on the (nz-1)(nr-1) unknowns as a sparse matrix
        (unknown (i, j) at (j-1)(nr-1) + i-1)."""
        import scipy.sparse as sp
        def tri(s, d, u):
            return sp.diags([s[1:], d, u[:-1]], [-1, 0, 1], format="csr")
        Lr = tri(*self.rows("r", which))
        Lz = tri(*self.rows("z"))
        return (sp.kron(sp.identity(self.nz - 1), Lr) + sp.kron(Lz, sp.identity(self.nr - 1))).tocsr()

    def solve_sparse(self, which: int, rhs) -> np.ndarray:
        """Direct sparse solve (small grids)."""
        import scipy.sparse.linalg as sla
        out = np.zeros(self.shape)
        x = sla.spsolve(self.matrix(which).tocsc(), rhs[1:-1, 1:-1].ravel())
        out[1:-1, 1:-1] = x.reshape(self.nz - 1, self.nr - 1)
        return out

    def solve(self, which: int, rhs) -> np.ndarray:
        """Separable exact solve: eigenbasis of the z rows, then one
        tridiagonal solve along r per z mode."""
        if self._solver is None:
            self._solver = {}
        if "z" not in self._solver:
            s, d, u = self.rows("z")
            Lz = np.diag(d) + np.diag(s[1:], -1) + np.diag(u[:-1], 1)
            lam, V = np.linalg.eig(Lz)
            o = np.argsort(lam.real)
            lam, V = lam.real[o], V.real[:, o]
            self._solver["z"] = (lam, V, np.linalg.inv(V))
        lam, V, Vi = self._solver["z"]
        if which not in self._solver:
            # Thomas factors along r, per z mode (they depend on the operator only)
            s, d, u = self.rows("r", which)
            n = d.size
            bp = np.empty((lam.size, n))
            fa = np.zeros((lam.size, n))
            bp[:, 0] = d[0] + lam
            for i in range(1, n):
                fa[:, i] = s[i] / bp[:, i - 1]
                bp[:, i] = d[i] + lam - fa[:, i] * u[i - 1]
            self._solver[which] = (fa, bp, u)
        fa, bp, u = self._solver[which]
        G = Vi @ rhs[1:-1, 1:-1]               # (modes, r)
        n = bp.shape[1]
        rp = np.empty_like(G)
        rp[:, 0] = G[:, 0]
        for i in range(1, n):
            rp[:, i] = G[:, i] - fa[:, i] * rp[:, i - 1]
        X = np.empty_like(G)
        X[:, n - 1] = rp[:, n - 1] / bp[:, n - 1]
        for i in range(n - 2, -1, -1):
            X[:, i] = (rp[:, i] - u[i] * X[:, i + 1]) / bp[:, i]
        out = np.zeros(self.shape)
        out[1:-1, 1:-1] = V @ X
        return out

    def residual(self, which: int, psi, rhs) -> float:
        """max |L psi - rhs| / max |rhs| over the inner nodes."""
        R = self.apply(which, psi)[1:-1, 1:-1] - rhs[1:-1, 1:-1]
        return float(np.abs(R).max() / max(np.abs(rhs[1:-1, 1:-1]).max(), 1e-300))

    def step(self):
        """One whole step: derivatives, DuFort-Frankel, right-hand sides,
        the two velocity solves, walls and the level rotation."""
        self.derivatives()
        self.explicit()
        self.rhs()
        self.f["u"][...] = self.solve(1, self.f["fu"])
        self.f["ww"][...] = self.solve(2, self.f["fw"])
        self.update()

    def energy(self) -> float:
        """Kinetic energy 1/2 sum (u^2 + v^2 + w^2) r dr dz with trapezoid
        weights on the nodes."""
        wr = np.zeros(self.nr + 1)
        wr[1:] += 0.5 * np.diff(self.r)
        wr[:-1] += 0.5 * np.diff(self.r)
        wz = np.zeros(self.nz + 1)
        wz[1:] += 0.5 * np.diff(self.z)
        wz[:-1] += 0.5 * np.diff(self.z)
        e = self.f["u"] ** 2 + self.f["ww"] ** 2 + self.f["v"] ** 2
        return float(0.5 * np.sum(wz[:, None] * wr[None, :] * self.r[None, :] * e))


def initial_state(m: Merid, seed: int, amp: float = 1e-4):
    """Temperature linear in r between the cylinders (0 inside, 1 outside)
    plus seeded U(-amp, amp) at the inner nodes; fluid at rest relative to
    the rotating walls (solid-body rotation, swirl 0 in the rotating frame,
    rotann2.f:81-82); previous levels equal to the current ones."""
    rng = np.random.default_rng(seed)
    r = m.r
    th = np.broadcast_to((r - r[0]) / (r[-1] - r[0]), m.shape).copy()
    th[1:-1, 1:-1] += rng.uniform(-amp, amp, size=(m.nz - 1, m.nr - 1))
    for k in FIELDS:
        m.f[k][...] = 0.0
    m.f["th"][...] = th
    m.f["thp"][...] = th
    m.f["thn"][...] = th


def time_steps(cfg_json: str, out_json: str):
    """CPU leg of the benchmark: a few whole oracle steps on one core."""
    import json
    try:
        from threadpoolctl import threadpool_limits
        lim = threadpool_limits(1)
    except Exception:  # pragma: no cover
        lim = None
    c = json.loads(cfg_json)
    m = Merid(c["nr"], c["nz"], c["dt"], c["Ra"], c["Re"], c["Pr"], r0=c.get("r0", 0.5), rmax=c.get("rmax", 1.0),
              H=c.get("H", 0.5), beta=c["beta"])
    initial_state(m, c["seed"])
    m.step()   # builds the z eigenbasis
    t0 = time.perf_counter()
    n = int(c["steps"])
    for _ in range(n):
        m.step()
    dt = time.perf_counter() - t0
    json.dump({"steps": n, "seconds": dt, "heat": m.heat().tolist()}, open(out_json, "w"))
    del lim


if __name__ == "__main__":
    time_steps(sys.argv[1], sys.argv[2])
