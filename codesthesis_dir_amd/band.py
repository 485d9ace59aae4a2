"""Batched tdma / ptdma / pdma on device arrays (include/vvband.h).

The reference solves one line per call (tdma, ptdma, pdma,
velvor2dperiodicrbconv.f:3046-3164); here a batch of lines is one launch with
the reference's arithmetic, so each solution is bit-for-bit the routine's.
Arrays are device tensors of shape (nsys, n) (row j of system s at [s, j]);
any strides are accepted, and the tensors' strides are passed to the library.
torch provides the device memory and the stream only.
"""
from __future__ import annotations

import ctypes

from ._lib import check, lib

TDMA, PTDMA, PDMA = 0, 1, 2
NOALLOC_WORDS = 8388608

_ARGT = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_longlong] + \
    [ctypes.c_void_p] * 7 + [ctypes.c_void_p]


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def solve(kind: int, a, b, c, r, out, d=None, e=None, stream=None, noalloc: bool = False):
    """Solves the systems of rows a, b, c (and d, e for PDMA) with right-hand
    sides r into out; all (nsys, n) float64 tensors on one device with the
    same strides.  stream: a torch.cuda.Stream (None: the current one).
    noalloc selects vv_band_solve_noalloc (no device allocation; batches up
    to NOALLOC_WORDS scratch words)."""
    import torch
    ts = [t for t in (a, b, c, d, e, r, out) if t is not None]
    for t in ts:
        if t.dtype != torch.float64 or not t.is_cuda or t.dim() != 2:
            raise ValueError("band.solve: float64 device tensors of shape (nsys, n)")
        if t.shape != r.shape or t.stride() != r.stride():
            raise ValueError("band.solve: every array needs the shape and strides of r")
    nsys, n = r.shape
    ss, sr = r.stride()
    if nsys == 1:
        ss = max(ss, 1)
    if n == 1:
        sr = max(sr, 1)
    st = stream if stream is not None else torch.cuda.current_stream(r.device)
    fn = lib().vv_band_solve_noalloc if noalloc else lib().vv_band_solve
    fn.argtypes = _ARGT
    fn.restype = ctypes.c_int
    with torch.cuda.device(r.device):
        code = fn(kind, n, nsys, ss, sr, _ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(e), _ptr(r), _ptr(out),
                  ctypes.c_void_p(st.cuda_stream))
    check(code, "vv_band_solve_noalloc" if noalloc else "vv_band_solve")
    return out
