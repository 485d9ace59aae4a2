"""Batched tdma / ptdma / pdma on device arrays (include/vvband.h).

The reference solves one line per call (tdma, ptdma, pdma,
velvor2dperiodicrbconv.f:3046-3164); here a batch of lines is one launch with
the reference's arithmetic, so each solution is bit-for-bit the routine's.

Device memory comes from the HIP runtime that libvvmi.so itself links
(DeviceArray below), so the buffers, the library and its streams share one
runtime instance in the process.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import check, lib

TDMA, PTDMA, PDMA = 0, 1, 2
NOALLOC_WORDS = 8388608
_H2D, _D2H = 1, 2

_ARGT = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong] + \
    [ctypes.c_void_p] * 8

_hip = None


def hip():
    """The libamdhip64 that libvvmi.so was linked against (already mapped)."""
    global _hip
    if _hip is None:
        lib()
        path = "libamdhip64.so.7"
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64.so" in line and "/torch/" not in line:
                    path = line.split()[-1]
                    break
        H = ctypes.CDLL(path)
        H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        H.hipFree.argtypes = [ctypes.c_void_p]
        H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        H.hipDeviceSynchronize.argtypes = []
        _hip = H
    return _hip


class DeviceArray:
    """A device allocation of the library's HIP runtime holding float64s."""

    def __init__(self, n: int):
        self.n = int(n)
        p = ctypes.c_void_p()
        if hip().hipMalloc(ctypes.byref(p), max(self.n, 1) * 8) != 0:
            raise MemoryError("hipMalloc failed")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(a.size)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size <= self.n
        if hip().hipMemcpy(self.ptr, a.ctypes.data, a.size * 8, _H2D) != 0:
            raise RuntimeError("hipMemcpy to the device failed")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float64)
        if hip().hipMemcpy(out.ctypes.data, self.ptr, self.n * 8, _D2H) != 0:
            raise RuntimeError("hipMemcpy from the device failed")
        return out

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _p(x):
    if x is None:
        return None
    return ctypes.c_void_p(x.ptr if isinstance(x, DeviceArray) else int(x))


def solve(kind: int, n: int, nsys: int, ss: int, sr: int, a, b, c, r, x, d=None, e=None,
          stream=None, noalloc: bool = False) -> int:
    """The C entry point on device arrays (DeviceArray or raw device
    addresses): row j of system s of each array at element s*ss + j*sr.
    stream: a hipStream_t address (None: the null stream).  noalloc selects
    vv_band_solve_noalloc.  Returns after the launch."""
    fn = lib().vv_band_solve_noalloc if noalloc else lib().vv_band_solve
    fn.argtypes = _ARGT
    fn.restype = ctypes.c_int
    code = fn(kind, n, nsys, ss, sr, _p(a), _p(b), _p(c), _p(d), _p(e), _p(r), _p(x),
              None if stream is None else ctypes.c_void_p(stream))
    return check(code, "vv_band_solve_noalloc" if noalloc else "vv_band_solve")


def solve_host(kind: int, a, b, c, r, d=None, e=None, interleaved: bool = True, pad: int = 0,
               noalloc: bool = False) -> np.ndarray:
    """Solves the (nsys, n) host systems on the GPU and returns x (nsys, n).
    interleaved stores the lines column-wise on the device (ss = 1, sr =
    nsys + pad), otherwise row-wise (ss = n + pad, sr = 1)."""
    r = np.asarray(r, dtype=np.float64)
    nsys, n = r.shape
    if interleaved:
        ss, sr, shape = 1, nsys + pad, (n, nsys + pad)
    else:
        ss, sr, shape = n + pad, 1, (nsys, n + pad)

    def lay(v):
        buf = np.zeros(shape)
        if interleaved:
            buf[:, :nsys] = np.asarray(v, dtype=np.float64).T
        else:
            buf[:, :n] = v
        return DeviceArray.from_numpy(buf)

    arrs = [None if v is None else lay(v) for v in (a, b, c, d, e, r)]
    x = DeviceArray(shape[0] * shape[1])
    solve(kind, n, nsys, ss, sr, arrs[0], arrs[1], arrs[2], arrs[5], x, d=arrs[3], e=arrs[4],
          noalloc=noalloc)
    hip().hipDeviceSynchronize()
    out = x.numpy().reshape(shape)
    for v in arrs + [x]:
        if v is not None:
            v.free()
    sol = out[:, :nsys].T if interleaved else out[:, :n]
    return np.ascontiguousarray(sol)
