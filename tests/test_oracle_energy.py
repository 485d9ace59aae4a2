"""The kinetic-energy integral of the oracle (vo_energy, v3_energy; the
reference has none, so the definition is the build's: one half of the sum
of the squared velocities over their Poisson unknowns, weighted by the cell
volume) pinned by closed forms: constant fields fix the index ranges and the
radius weights, single points fix the excluded walls and the duplicate
column k = np, and solid-body rotation fixes it as a volume integral."""
import numpy as np
import pytest

import oracle as O
import oracle3 as O3


def _box(geom, nr=12, np_=16):
    if geom:
        return O.Run(nr, np_, 1e-5, 1e4, 1.0, geom=1, r0=0.5, Re=10.0)
    return O.Run(nr, np_, 1e-5, 1e4, 1.0, width=1.0)


@pytest.mark.parametrize("geom", [0, 1])
def test_2d_constant_fields(geom):
    run = _box(geom)
    nr, np_, dr, dp, r0 = run.nr, run.np, run.p.dr, run.p.dp, 0.5
    U, V = 0.7, -1.3
    run["uel"][...] = U
    run["vel"][...] = V
    if geom:
        sv = nr * r0 + dr * nr * nr / 2.0             # sum_{i=1..nr} (r0 + (i-1/2) dr)
        su = (nr - 1) * r0 + dr * (nr - 1) * nr / 2.0  # sum_{i=1..nr-1} (r0 + i dr)
    else:
        sv, su = nr, nr - 1
    want = 0.5 * dr * dp * np_ * (V * V * sv + U * U * su)
    assert abs(run.energy() - want) <= 1e-14 * want


@pytest.mark.parametrize("geom", [0, 1])
def test_2d_single_points(geom):
    run = _box(geom)
    nr, np_, dr, dp = run.nr, run.np, run.p.dr, run.p.dp
    run["vel"][5, 3] = 2.0        # column k=5, row i=3
    w = (0.5 + 2.5 * dr) if geom else 1.0
    assert run.energy() == pytest.approx(0.5 * dr * dp * w * 4.0, rel=1e-15)
    run["vel"][...] = 0.0
    run["vel"][:, 0] = run["vel"][:, nr + 1] = 1.0   # ghost cells
    run["uel"][:, 0] = run["uel"][:, nr] = 1.0       # walls
    run["uel"][np_, :] = run["vel"][np_, :] = 1.0    # the duplicate column k = np
    assert run.energy() == 0.0


def test_2d_solid_body_rotation():
    """v = Omega r in the annulus: 1/2 int Omega^2 r^2 r dr dphi =
    pi Omega^2 (rmax^4 - r0^4) / 4; the midpoint sum converges as dr^2."""
    errs = []
    for nr in (32, 64):
        run = O.Run(nr, 64, 1e-5, 1e4, 1.0, geom=1, r0=0.5, Re=10.0)
        r = 0.5 + (np.arange(run.ld) - 0.5) * run.p.dr
        run["vel"][...] = 3.0 * r[None, :]
        want = np.pi * 9.0 * (1.0 - 0.5 ** 4) / 4.0
        errs.append(abs(run.energy() - want) / want)
    assert errs[0] < 1e-3 and 3.5 < errs[0] / errs[1] < 4.5, errs


def test_3d_constant_fields():
    nr, nz, np_ = 8, 6, 10
    run = O3.Run3(nr, nz, np_, 1e-4, 1e4, 30.0, 1.0, 0.0)
    dr, dz, dp, r0 = run.p.dr, run.p.dz, run.p.dp, 0.5
    U, V, W = 0.4, -0.9, 1.7
    run["uel"] = U
    run["vel"] = V
    run["wel"] = W
    sc = nr * r0 + dr * nr * nr / 2.0
    sf = (nr - 1) * r0 + dr * (nr - 1) * nr / 2.0
    want = 0.5 * dr * dz * dp * np_ * (U * U * nz * sf + V * V * nz * sc + W * W * (nz - 1) * sc)
    assert abs(run.energy() - want) <= 1e-14 * want


def test_3d_walls_and_duplicate_plane_excluded():
    nr, nz, np_ = 8, 6, 10
    run = O3.Run3(nr, nz, np_, 1e-4, 1e4, 30.0, 1.0, 0.0)
    for n in ("uel", "vel", "wel"):
        a = run[n]
        a[np_] = 1.0
        a[:, 0, :] = a[:, nz + 1, :] = 1.0
        a[:, :, 0] = a[:, :, nr + 1] = 1.0
    run["uel"][:, :, nr] = 1.0      # outer wall of u (face nr)
    run["wel"][:, nz, :] = 1.0      # top wall of w (face nz)
    assert run.energy() == 0.0


def test_3d_solid_body_rotation():
    """v = Omega r: pi Omega^2 zmax (rmax^4 - r0^4) / 4, second order."""
    errs = []
    for nr in (16, 32):
        run = O3.Run3(nr, 4, 16, 1e-4, 1e4, 30.0, 1.0, 0.0)
        r = 0.5 + (np.arange(nr + 2) - 0.5) * run.p.dr
        run["vel"] = 2.0 * r[None, None, :]
        want = np.pi * 4.0 * 0.5 * (1.0 - 0.5 ** 4) / 4.0
        errs.append(abs(run.energy() - want) / want)
    assert errs[0] < 1e-3 and 3.5 < errs[0] / errs[1] < 4.5, errs
