"""Host interface of the meridional-plane step with swirl on stretched nodes
on the HIP library (include/vvmerid.h): one object per run, methods named
after the routines of the step (derivatives, explicit, rhs, poisson, update,
heat), fields as numpy arrays of shape (nz+1, nr+1) (X[j, i] is X(r_i, z_j)).

The compute path is libvvmi.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib

_P = ctypes.POINTER(ctypes.c_double)
FIELDS = ["w", "wp", "wn", "v", "vp", "vn", "th", "thp", "thn", "u", "ww", "fu", "fw"]
FIELD_IDS = {n: k for k, n in enumerate(FIELDS)}
KERNELS = ["deriv", "explicit", "rhs", "zfwd", "rsolve", "zinv", "update"]
# the work of these kernels is counted in floating-point operations, of the others in bytes
FLOP_KERNELS = ("zfwd", "zinv")


class VVMConfig(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nz", ctypes.c_int), ("dt", ctypes.c_double), ("Ra", ctypes.c_double),
                ("Re", ctypes.c_double), ("Pr", ctypes.c_double)]


def tanh_nodes(n: int, a: float, b: float, beta: float) -> np.ndarray:
    """n+1 nodes on [a, b] clustered toward both ends:
    a + (b - a) (1 + tanh(beta (2 q/n - 1)) / tanh(beta)) / 2, q = 0..n
    (beta = 0: uniform); the ends are exactly a and b."""
    xi = np.arange(n + 1, dtype=np.float64) / n
    s = xi if beta == 0 else 0.5 * (1.0 + np.tanh(beta * (2.0 * xi - 1.0)) / np.tanh(beta))
    x = a + (b - a) * s
    x[0], x[-1] = a, b
    return x


@dataclass
class MeridConfig:
    """A meridional plane of nr x nz intervals between the cylinders r0 and
    rmax, of height H, nodes tanh-clustered toward the walls (beta);
    Re = sqrt(Ta)."""
    nr: int
    nz: int
    dt: float
    Ra: float
    Re: float
    Pr: float
    r0: float = 0.5
    rmax: float = 1.0
    H: float = 0.5
    beta: float = 1.5

    def c(self) -> VVMConfig:
        return VVMConfig(self.nr, self.nz, self.dt, self.Ra, self.Re, self.Pr)

    def nodes(self):
        return tanh_nodes(self.nr, self.r0, self.rmax, self.beta), tanh_nodes(self.nz, 0.0, self.H, self.beta)


def _bind():
    L = lib()
    if getattr(L, "_vvm_bound", False):
        return L
    vp = ctypes.c_void_p
    L.vvm_create.argtypes = [ctypes.POINTER(VVMConfig), _P, _P, ctypes.c_int, ctypes.POINTER(vp)]
    for fn in ("vvm_destroy", "vvm_derivatives", "vvm_explicit", "vvm_rhs", "vvm_update"):
        getattr(L, fn).argtypes = [vp]
    L.vvm_put.argtypes = [vp, ctypes.c_int, _P]
    L.vvm_get.argtypes = [vp, ctypes.c_int, _P]
    L.vvm_get_derivative.argtypes = [vp, ctypes.c_int, _P]
    L.vvm_poisson.argtypes = [vp, ctypes.c_int]
    L.vvm_residual.argtypes = [vp, ctypes.c_int, _P]
    L.vvm_step.argtypes = [vp, ctypes.c_int]
    L.vvm_heat.argtypes = [vp, _P]
    L.vvm_energy.argtypes = [vp, _P]
    L.vvm_profile.argtypes = [vp, ctypes.c_int]
    L.vvm_graphs.argtypes = [vp, ctypes.c_int]
    L.vvm_profile_read.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _P, _P]
    L._vvm_bound = True
    return L


class Merid:
    """One meridional-plane run resident on a GPU.  r, z: the nodes (by
    default the config's tanh nodes)."""

    def __init__(self, cfg: MeridConfig, device: int = 0, r=None, z=None):
        self.cfg = cfg
        self._lib = _bind()
        rn, zn = cfg.nodes()
        self.r = np.ascontiguousarray(rn if r is None else r, dtype=np.float64)
        self.z = np.ascontiguousarray(zn if z is None else z, dtype=np.float64)
        if self.r.size != cfg.nr + 1 or self.z.size != cfg.nz + 1:
            raise ValueError("node counts must be nr+1 and nz+1")
        h = ctypes.c_void_p()
        c = cfg.c()
        check(self._lib.vvm_create(ctypes.byref(c), self.r.ctypes.data_as(_P), self.z.ctypes.data_as(_P), device,
                                   ctypes.byref(h)), "vvm_create")
        self._h = h
        self.shape = (cfg.nz + 1, cfg.nr + 1)

    def close(self):
        if self._h:
            self._lib.vvm_destroy(self._h)
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
        check(self._lib.vvm_put(self._h, FIELD_IDS[name], a.ctypes.data_as(_P)), "vvm_put")

    def get(self, name: str) -> np.ndarray:
        out = np.empty(self.shape)
        check(self._lib.vvm_get(self._h, FIELD_IDS[name], out.ctypes.data_as(_P)), "vvm_get")
        return out

    def derivative(self, d: int) -> np.ndarray:
        out = np.empty(self.shape)
        check(self._lib.vvm_get_derivative(self._h, d, out.ctypes.data_as(_P)), "vvm_get_derivative")
        return out

    def derivatives(self):
        check(self._lib.vvm_derivatives(self._h), "vvm_derivatives")

    def explicit(self):
        check(self._lib.vvm_explicit(self._h), "vvm_explicit")

    def rhs(self):
        check(self._lib.vvm_rhs(self._h), "vvm_rhs")

    def poisson(self, which: int):
        check(self._lib.vvm_poisson(self._h, which), "vvm_poisson")

    def residual(self, which: int) -> float:
        x = ctypes.c_double()
        check(self._lib.vvm_residual(self._h, which, ctypes.byref(x)), "vvm_residual")
        return x.value

    def update(self):
        check(self._lib.vvm_update(self._h), "vvm_update")

    def step(self, n: int = 1):
        check(self._lib.vvm_step(self._h, n), "vvm_step")

    def heat(self):
        out = np.zeros(2)
        check(self._lib.vvm_heat(self._h, out.ctypes.data_as(_P)), "vvm_heat")
        return out

    def energy(self) -> float:
        e = ctypes.c_double()
        check(self._lib.vvm_energy(self._h, ctypes.byref(e)), "vvm_energy")
        return e.value

    def graphs(self, enable: bool):
        """Replay whole steps from HIP graphs (default) or launch them one
        by one."""
        check(self._lib.vvm_graphs(self._h, int(enable)), "vvm_graphs")

    def profile(self, enable: bool):
        check(self._lib.vvm_profile(self._h, int(enable)), "vvm_profile")

    def profile_read(self):
        out = {}
        for k, name in enumerate(KERNELS):
            n, ms, b = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
            check(self._lib.vvm_profile_read(self._h, k, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)),
                  "vvm_profile_read")
            unit = "flops" if name in FLOP_KERNELS else "bytes"
            out[name] = {"launches": n.value, "ms": ms.value, unit: b.value}
        return out
