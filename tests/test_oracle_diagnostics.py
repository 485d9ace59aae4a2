"""The 3D diagnostics of the oracle pinned by mathematics: the local
Nusselt numbers at mid-height (sect360v2mg.f:318-325) equal 1 for the
conduction profile to third order in dr (the one-sided weights
(-184, 225, -50, 9)/60 are the third-order derivative at the wall of the
staggered points); the Fourier-mode integrals (after rot3dmghor.f:954-991)
of a single azimuthal mode pick that mode alone, with the closed-form
trapezoid weights."""
import numpy as np
import pytest

import oracle3 as O3


def _conduction(nr, nz, np_, r0=0.5, rmax=1.0, hot_inside=False):
    """Conduction between the cylinders: T = 1 - ln(r/rmax)/ln(r0/rmax)
    (0 inside, 1 outside, as rot3dmg.f:120), or 1 - that."""
    run = O3.Run3(nr, nz, np_, 1e-4, 1e4, 30.0, 1.0, 0.0, r0=r0, rmax=rmax)
    dr = (rmax - r0) / nr
    i = np.arange(nr + 2)
    r = np.where(i == 0, r0, np.where(i == nr + 1, rmax, r0 + (i - 0.5) * dr))
    t = 1.0 - np.log(r / rmax) / np.log(r0 / rmax)
    run["th"] = np.broadcast_to(1.0 - t if hot_inside else t, run.shape)
    return run


@pytest.mark.parametrize("hot_inside,nu", [(False, 1.0), (True, -1.0)])
def test_local_nu_of_conduction_is_one_to_third_order(hot_inside, nu):
    """Nu = -r ln(r0) dT/dr at rmax = 1 is +1 for the cold inner wall of
    rot3dmg.f and -1 for the hot one of rot3dmghor.f, at both cylinders."""
    errs = []
    for nr in (16, 32):
        a, b = _conduction(nr, 8, 12, hot_inside=hot_inside).local_nu()
        errs.append(max(np.max(np.abs(a - nu)), np.max(np.abs(b - nu))))
    assert errs[0] < 1e-3 and 6.0 < errs[0] / errs[1] < 10.0, errs


def test_local_nu_reads_mid_height_only():
    run = _conduction(16, 8, 12)
    base = run.local_nu()
    run["th"][:, 3, 1] += 1.0       # j = 3 != nz/2 = 4
    assert all(np.array_equal(x, y) for x, y in zip(run.local_nu(), base))
    run["th"][5, 4, 1] += 1.0       # j = nz/2, plane k = 5, row i = 1
    a, b = run.local_nu()
    assert np.array_equal(np.nonzero(a - base[0])[0], [5]) and np.array_equal(b, base[1])


@pytest.mark.parametrize("m,kind", [(1, "cos"), (3, "sin"), (5, "cos")])
def test_modes_pick_a_single_mode(m, kind):
    nr, nz, np_ = 10, 6, 24
    run = O3.Run3(nr, nz, np_, 1e-4, 1e4, 30.0, 1.0, 0.0)
    dp, dr, dz, r0 = run.p.dp, run.p.dr, run.p.dz, 0.5
    k = np.arange(np_ + 1)
    ph = (np.cos if kind == "cos" else np.sin)(m * k * dp)
    run["th"] = np.broadcast_to(ph[:, None, None], run.shape)
    c, s = run.modes(6)
    weight = ((nr + 1) * r0 + dr * (nr + 1) ** 2 / 2.0) * (nz + 1)   # sum f_i f_j r_i
    want = 0.5 * np_ * dp * dr * dz * weight
    hit, other = (c, s) if kind == "cos" else (s, c)
    assert abs(hit[m - 1] - want) <= 1e-13 * want
    assert np.max(np.abs(np.delete(hit, m - 1))) <= 1e-13 * want
    assert np.max(np.abs(other)) <= 1e-13 * want
