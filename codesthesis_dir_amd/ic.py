"""Seeded initial conditions of the benchmark and test workloads.

The temperature perturbation is a splitmix64 stream (counter-based, so it
is generated vectorised): draw j gives U[0,1) from the top 53 bits of
mix(seed + (j+1)*0x9E3779B97F4A7C15).  Draws run over k = 0..np-1 (outer)
and i = 1..nr (inner); column k = np repeats column 0.
"""
from __future__ import annotations

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)


def splitmix64_uniform(seed: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        j = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + j * _G
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def box_temperature(nr: int, np_: int, r0: float, rmax: float, amp: float, seed: int,
                    ld: int | None = None) -> np.ndarray:
    """Linear conduction profile between T(r0)=0 and T(rmax)=1 on the cell
    rows plus U(-amp, amp) noise; wall rows 0 and 1 (rb.f:139-188 layout)."""
    ld = ld if ld is not None else nr + 2
    dr = (rmax - r0) / nr
    th = np.zeros((np_ + 1, ld))
    u = splitmix64_uniform(seed, nr * np_).reshape(np_, nr)
    r = r0 + (np.arange(1, nr + 1) - 0.5) * dr
    th[:np_, 1:nr + 1] = (r - r0)[None, :] / (rmax - r0) + amp * (2.0 * u - 1.0)
    th[np_] = th[0]
    th[:, 0] = 0.0
    th[:, nr + 1] = 1.0
    return th


def merid_temperature(r: np.ndarray, nz: int, amp: float, seed: int) -> np.ndarray:
    """Meridional plane (nodes r[0..nr], z rows 0..nz): temperature linear in
    r between the cylinders (0 inside, 1 outside) plus U(-amp, amp) at the
    inner nodes, drawn by numpy's default_rng(seed) as one (nz-1) x (nr-1)
    block, j outer."""
    nr = r.size - 1
    th = np.broadcast_to((r - r[0]) / (r[-1] - r[0]), (nz + 1, nr + 1)).copy()
    rng = np.random.default_rng(seed)
    th[1:-1, 1:-1] += rng.uniform(-amp, amp, size=(nz - 1, nr - 1))
    return th
