"""Python binding of the request loop (include/vvserve.h): many runs served
on one GPU at once, admitted by slot and device footprint, each on its own
stream and worker thread inside libvvmi.so.

    with Server(slots=8, mem_budget=32e9) as srv:
        rid = srv.submit(cfg, nsteps=2000, th0=th0)
        res = srv.wait()
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .velvor import BoxConfig

_P = ctypes.POINTER(ctypes.c_double)


class VVRequest(ctypes.Structure):
    _fields_ = [("cfg", L.VVConfig), ("nsteps", ctypes.c_int), ("ld", ctypes.c_int), ("th0", _P),
                ("th_out", _P), ("vel_out", _P), ("user", ctypes.c_void_p)]


class VVResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_longlong), ("status", ctypes.c_int), ("steps", ctypes.c_int),
                ("trans1", ctypes.c_double), ("trans2", ctypes.c_double), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double), ("user", ctypes.c_void_p),
                ("error", ctypes.c_char * 128)]


class VVServerStats(ctypes.Structure):
    _fields_ = [("queued", ctypes.c_int), ("running", ctypes.c_int), ("done", ctypes.c_int),
                ("peak_running", ctypes.c_int), ("mem_in_use", ctypes.c_double), ("peak_mem", ctypes.c_double),
                ("resident", ctypes.c_int), ("reused", ctypes.c_int), ("created", ctypes.c_int)]


def _bind():
    lib = L.lib()
    if getattr(lib, "_serve_bound", False):
        return lib
    vp = ctypes.c_void_p
    lib.vv_run_footprint.restype = ctypes.c_double
    lib.vv_run_footprint.argtypes = [ctypes.POINTER(L.VVConfig)]
    lib.vv_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.POINTER(vp)]
    lib.vv_server_submit.restype = ctypes.c_longlong
    lib.vv_server_submit.argtypes = [vp, ctypes.POINTER(VVRequest)]
    lib.vv_server_wait.argtypes = [vp, ctypes.POINTER(VVResult), ctypes.c_double]
    lib.vv_server_stats_read.argtypes = [vp, ctypes.POINTER(VVServerStats)]
    lib.vv_server_destroy.argtypes = [vp]
    lib._serve_bound = True
    return lib


def footprint(cfg: BoxConfig) -> float:
    c = cfg.c()
    return _bind().vv_run_footprint(ctypes.byref(c))


class Server:
    """The request loop.  submit() returns the request id; the output arrays
    (final temperature and v velocity, shape (np+1, nr+2)) are kept here and
    handed back with the result by wait()."""

    def __init__(self, slots: int = 8, mem_budget: float = 32e9, device: int = 0):
        self._lib = _bind()
        self._h = ctypes.c_void_p()
        L.check(self._lib.vv_server_create(device, slots, mem_budget, ctypes.byref(self._h)), "vv_server_create")
        self._keep = {}

    def submit(self, cfg: BoxConfig, nsteps: int, th0: np.ndarray | None = None, outputs: bool = False) -> int:
        ld = cfg.nr + 2
        t = None
        if th0 is not None:
            t = np.ascontiguousarray(th0[:, :ld], dtype=np.float64)
            if t.shape != (cfg.np + 1, ld):
                raise ValueError(f"th0: shape {t.shape}, expected {(cfg.np + 1, ld)}")
        th_out = np.zeros((cfg.np + 1, ld)) if outputs else None
        vel_out = np.zeros((cfg.np + 1, ld)) if outputs else None
        req = VVRequest(cfg.c(), nsteps, ld,
                        t.ctypes.data_as(_P) if t is not None else None,
                        th_out.ctypes.data_as(_P) if outputs else None,
                        vel_out.ctypes.data_as(_P) if outputs else None, None)
        rid = self._lib.vv_server_submit(self._h, ctypes.byref(req))
        if rid < 0:
            raise L.VVError(int(rid), "vv_server_submit")
        self._keep[rid] = (th_out, vel_out)
        return rid

    def wait(self, timeout: float = -1.0):
        """The next completed result as a dict, or None after `timeout` s."""
        r = VVResult()
        st = self._lib.vv_server_wait(self._h, ctypes.byref(r), timeout)
        if st == 1:
            return None
        L.check(st, "vv_server_wait")
        th_out, vel_out = self._keep.pop(r.id, (None, None))
        return {"id": r.id, "status": r.status, "steps": r.steps, "trans1": r.trans1, "trans2": r.trans2,
                "t_submit": r.t_submit, "t_admit": r.t_admit, "t_done": r.t_done, "th": th_out, "vel": vel_out,
                "error": r.error.decode(errors="replace")}

    def stats(self) -> dict:
        s = VVServerStats()
        L.check(self._lib.vv_server_stats_read(self._h, ctypes.byref(s)), "vv_server_stats_read")
        return {"queued": s.queued, "running": s.running, "done": s.done, "peak_running": s.peak_running,
                "mem_in_use": s.mem_in_use, "peak_mem": s.peak_mem, "resident": s.resident, "reused": s.reused,
                "created": s.created}

    def close(self):
        if self._h:
            self._lib.vv_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
