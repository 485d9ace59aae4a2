"""ctypes binding of libvvmi.so (the C ABI declared in include/vvmi.h).

The product path is the HIP library only: if libvvmi.so is missing this
module raises instead of falling back to any CPU code.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(PKG, "libvvmi.so")

VV_OK, VV_NOT_CONVERGED, VV_EINVAL, VV_EHIP, VV_ENOMEM, VV_ENONFINITE = 0, 1, -1, -2, -3, -4

(VV_C, VV_CP, VV_CN, VV_TH, VV_THP, VV_THN, VV_UEL, VV_VEL, VV_UELP, VV_VELP,
 VV_VRH, VV_URH, VV_MEAN_T2, VV_MEAN_T, VV_MEAN_U2, VV_MEAN_V2) = range(16)

FIELD_IDS = {"C": VV_C, "Cp": VV_CP, "Cn": VV_CN, "th": VV_TH, "thp": VV_THP, "thn": VV_THN,
             "uel": VV_UEL, "vel": VV_VEL, "uelp": VV_UELP, "velp": VV_VELP, "vrh": VV_VRH,
             "urh": VV_URH, "mean_T2": VV_MEAN_T2, "mean_T": VV_MEAN_T, "mean_U2": VV_MEAN_U2,
             "mean_V2": VV_MEAN_V2}


class VVConfig(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("np", ctypes.c_int), ("dt", ctypes.c_double),
                ("r0", ctypes.c_double), ("rmax", ctypes.c_double), ("width", ctypes.c_double),
                ("Ra", ctypes.c_double), ("Pr", ctypes.c_double), ("robert", ctypes.c_double),
                ("rtol", ctypes.c_double), ("max_cycles", ctypes.c_int), ("nu1", ctypes.c_int),
                ("nu2", ctypes.c_int), ("stats", ctypes.c_int), ("geom", ctypes.c_int),
                ("Re", ctypes.c_double), ("direct_cols", ctypes.c_int), ("precision", ctypes.c_int),
                ("lines", ctypes.c_int)]

VV_GEOM_BOX, VV_GEOM_ANNULUS = 0, 1


class VVStepInfo(ctypes.Structure):
    _fields_ = [("trans1", ctypes.c_double), ("trans2", ctypes.c_double),
                ("cflu", ctypes.c_double), ("cflv", ctypes.c_double),
                ("relres_v", ctypes.c_double), ("relres_u", ctypes.c_double),
                ("cycles_v", ctypes.c_int), ("cycles_u", ctypes.c_int),
                ("nclamp", ctypes.c_int), ("status", ctypes.c_int)]


_lib = None


def lib():
    """Load libvvmi.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with __graft_entry__.build() "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIBPATH)
        L.vv_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


class VVError(RuntimeError):
    def __init__(self, code, where):
        msg = lib().vv_last_error().decode(errors="replace")
        super().__init__(f"{where}: status {code}: {msg}")
        self.code = code


def check(code, where):
    if code < 0:
        raise VVError(code, where)
    return code
