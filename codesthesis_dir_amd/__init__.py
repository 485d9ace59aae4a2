"""MI355X-native velocity-vorticity time step (HIP for gfx950).

The compute path is libvvmi.so (C ABI in include/vvmi.h); this package holds
its build recipe and a thin host mirror of the reference routines.
"""
from .velvor import AnnulusConfig, BoxConfig, PeriodicBox  # noqa: F401

__all__ = ["AnnulusConfig", "BoxConfig", "PeriodicBox"]
