"""The barrier sector of sect360v2mg.f on the GPU (include/vvsector.h)
against its oracle (oracle/vvs_oracle.c, pinned in
tests/test_oracle_sector.py):

* the explicit routines, the heat flux, the right-hand sides and the update
  evaluate the oracle's expressions in its order: bit for bit, with the
  hybrid upwind switch on in part of the cells;
* the solves: within 1e-11 of the oracle's dense solve, psip and the walls
  exactly, the mirrored ghosts as the reference writes them;
* whole steps from the reference's initial state: within 1e-10 of the
  oracle's;
* at 129 x 129 x 257 (nr = nz = 128, np = 256), where the dense solves are
  out of reach: after 20 steps every solve satisfies the reference's
  residual operator to rounding and the explicit routines of the next step
  equal the oracle's bit for bit.
"""
import os
import sys

import numpy as np
import pytest

import oracle_sector as OS

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from bench import sector_state  # noqa: E402

pytestmark = pytest.mark.gpu

PHYS = dict(dt=1e-5, Ra=1e5, Re=100.0, Ptl=7.0, Gr=10.0)
NAMES = ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "uel", "vel", "wel", "An", "Bn", "Cn", "thn",
         "uelp", "velp", "welp", "urh", "vrh", "wrh")


def _gpu(nr, nz, np_, phys=PHYS):
    from codesthesis_dir_amd.sector import Sector, SectorConfig
    return Sector(SectorConfig(nr=nr, nz=nz, np=np_, **phys))


def _pair(nr, nz, np_, seed, scale=1.0, phys=PHYS):
    run = _gpu(nr, nz, np_, phys)
    ora = OS.Sector(nr, nz, np_, **phys)
    rng = np.random.default_rng(seed)
    for n in NAMES:
        a = scale * rng.standard_normal(ora.shape)
        ora[n] = a
        run.put(n, a)
    return run, ora


@pytest.mark.parametrize("shape", [(6, 6, 8), (9, 7, 13), (16, 12, 24)])
def test_explicit_routines_bitexact(shape):
    nr, nz, np_ = shape
    print(f"[gpu] sector explicit {shape}", flush=True)
    run, ora = _pair(nr, nz, np_, seed=nr + 3 * nz + 7 * np_)
    run.vorticity(0)
    run.temperature()
    for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
        ora.call(rt)
    for n in ("An", "Bn", "Cn", "thn"):
        assert np.array_equal(run.get(n), ora[n]), n
    assert run.heat() == ora.heat()
    for which, rt, rhs in ((1, "rhsw", "wrh"), (2, "rhsv", "vrh"), (3, "rhsu", "urh")):
        run.rhs(which)
        ora.call(rt)
        assert np.array_equal(run.get(rhs), ora[rhs]), rhs
    run.update()
    ora.call("update")
    for n in ("A", "Ap", "An", "B", "Bp", "Bn", "C", "Cp", "Cn", "th", "thp", "thn"):
        assert np.array_equal(run.get(n), ora[n]), n
    run.close()


def test_single_routines_match_the_fused_launch():
    """vvs_vorticity(1/2/3) run one routine each: the same values as the
    three in one launch"""
    nr, nz, np_ = 9, 7, 13
    print("[gpu] sector single routines", flush=True)
    run, ora = _pair(nr, nz, np_, seed=2)
    for which, rt, n in ((1, "avorticity", "An"), (2, "bvorticity", "Bn"), (3, "cvorticity", "Cn")):
        run.vorticity(which)
        ora.call(rt)
        assert np.array_equal(run.get(n), ora[n]), n
    run.close()


@pytest.mark.parametrize("which", [1, 2, 3])
@pytest.mark.parametrize("shape", [(6, 5, 8), (9, 7, 12), (12, 6, 9), (8, 6, 32)])
def test_poisson_matches_dense_solve(which, shape):
    """np = 8, 32: the phi transforms by FFT; 12, 9: by GEMM."""
    nr, nz, np_ = shape
    print(f"[gpu] sector poisson {which} {shape}", flush=True)
    run, ora = _pair(nr, nz, np_, seed=which + nr)
    psi, psip, rhs = OS.PSI[which]
    old = ora[psi].copy()
    run.poisson(which)
    ora.poisson(which)
    got, want = run.get(psi), ora[psi]
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))
    # zeros where the oracle writes zeros, untouched points untouched
    assert np.array_equal(got == 0.0, want == 0.0)
    untouched = want == old
    assert np.array_equal(got[untouched], old[untouched])
    assert np.array_equal(run.get(psip), ora[psip])
    m, _ = ora.residual(which, got, ora[rhs])
    assert m <= 1e-10 * np.max(np.abs(ora[rhs])) * max(nr, nz, np_) ** 2
    run.close()


@pytest.mark.parametrize("which", [1, 2, 3])
@pytest.mark.parametrize("shape", [(4, 4, 2048), (6, 5, 1024), (5, 4, 4), (7, 5, 5)])
def test_poisson_residual_at_size_limits(which, shape):
    """Where the dense solve is out of reach (np = 1024, 2048: the FFT path
    at its largest sizes) and at the smallest np (4, 5: the GEMM path): the
    reference's residual operator vanishes to rounding, psip is the old psi
    and the barrier planes are as the reference writes them."""
    nr, nz, np_ = shape
    print(f"[gpu] sector poisson limits {which} {shape}", flush=True)
    run, ora = _pair(nr, nz, np_, seed=17 + which + np_)
    psi, psip, rhs = OS.PSI[which]
    old = ora[psi].copy()
    run.poisson(which)
    got = run.get(psi)
    m, _ = ora.residual(which, got, ora[rhs])
    h = min(ora.p.dr, ora.p.dz, ora.p.r0 * ora.p.dp)
    scale = np.max(np.abs(ora[rhs])) + np.max(np.abs(got)) * 8.0 / h ** 2
    assert m <= 1e-13 * scale, (m, scale)
    ie, je, ke = {1: (nr + 1, nz, np_ + 1), 2: (nr + 1, nz + 1, np_), 3: (nr, nz + 1, np_ + 1)}[which]
    assert np.array_equal(run.get(psip)[:ke + 1, :je + 1, :ie + 1], old[:ke + 1, :je + 1, :ie + 1])
    if which == 2:
        assert np.all(got[0] == 0.0) and np.all(got[np_] == 0.0)
    else:
        i1 = nr if which == 1 else nr - 1
        j1 = nz - 1 if which == 1 else nz
        assert np.array_equal(got[0, 1:j1 + 1, 1:i1 + 1], -got[1, 1:j1 + 1, 1:i1 + 1])
        assert np.array_equal(got[np_ + 1, 1:j1 + 1, 1:i1 + 1], -got[np_, 1:j1 + 1, 1:i1 + 1])
    run.close()


def test_create_rejects_out_of_range_sizes():
    from codesthesis_dir_amd._lib import VVError
    from codesthesis_dir_amd.sector import Sector, SectorConfig
    for nr, nz, np_ in ((3, 6, 8), (6, 3, 8), (6, 6, 3), (6, 6, 2049), (6, 4097, 8)):
        with pytest.raises(VVError):
            Sector(SectorConfig(nr=nr, nz=nz, np=np_, **PHYS))


def _oracle_run(nr, nz, np_, phys):
    ora = OS.Sector(nr, nz, np_, **phys)
    for k, v in sector_state(nr, nz, np_).items():
        ora[k] = v
    return ora


def test_trajectory_matches_oracle():
    nr, nz, np_ = 10, 8, 12
    print("[gpu] sector trajectory", flush=True)
    run = _gpu(nr, nz, np_)
    ora = _oracle_run(nr, nz, np_, PHYS)
    for k, v in sector_state(nr, nz, np_).items():
        run.put(k, v)
    for s in range(10):
        tg = run.step(1)
        to = ora.step()
        assert abs(tg - to) <= 1e-10 * abs(to), (s, tg, to)
    for n in ("A", "B", "C", "th", "uel", "vel", "wel", "Ap", "thp", "welp"):
        got, want = run.get(n), ora[n]
        scale = max(np.max(np.abs(want)), 1e-300)
        assert np.max(np.abs(got - want)) <= 1e-10 * scale, n
    run.close()


def test_sector_129x129x257_grid_properties():
    """rot3dmg.f's rotating-annulus physics (Ra = 1e7, Ta = 1e6, Pr = 7) in
    the sector at nr = nz = 128, np = 256 from the reference's initial state:
    20 steps, then the next step's explicit routines against the oracle and
    its solves through the reference's residual operators."""
    nr, nz, np_ = 128, 128, 256
    phys = dict(dt=1e-6, Ra=1e7, Re=1000.0, Ptl=7.0, Gr=0.0)
    print(f"[gpu] sector full grid {nr}x{nz}x{np_}", flush=True)
    run = _gpu(nr, nz, np_, phys)
    for k, v in sector_state(nr, nz, np_).items():
        run.put(k, v)
    t = run.step(20)
    assert np.isfinite(t)
    ora = OS.Sector(nr, nz, np_, **phys)
    for n in NAMES:
        a = run.get(n)
        assert np.all(np.isfinite(a)), n
        ora[n] = a
    th = ora["th"]
    assert th.min() > -0.25 and th.max() < 1.01
    run.vorticity(0)
    run.temperature()
    for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
        ora.call(rt)
    for n in ("An", "Bn", "Cn", "thn"):
        assert np.array_equal(run.get(n), ora[n]), n
    assert run.heat() == ora.heat()
    for which, rt in ((1, "rhsw"), (2, "rhsv"), (3, "rhsu")):
        psi, psip, rhs = OS.PSI[which]
        run.rhs(which)
        ora.call(rt)
        f = run.get(rhs)
        assert np.array_equal(f, ora[rhs]), rhs
        old = run.get(psi)
        run.poisson(which)
        got = run.get(psi)
        m, _ = ora.residual(which, got, f)
        h = min(ora.p.dr, ora.p.dz, ora.p.r0 * ora.p.dp)
        scale = np.max(np.abs(f)) + np.max(np.abs(got)) * 8.0 / h ** 2
        assert m < 1e-13 * scale, (which, m, scale)
        pp = run.get(psip)
        ie, je, ke = {1: (nr + 1, nz, np_ + 1), 2: (nr + 1, nz + 1, np_), 3: (nr, nz + 1, np_ + 1)}[which]
        assert np.array_equal(pp[:ke + 1, :je + 1, :ie + 1], old[:ke + 1, :je + 1, :ie + 1])
        if which == 2:
            assert np.all(got[0] == 0.0) and np.all(got[np_] == 0.0)
        else:
            i1 = nr if which == 1 else nr - 1
            j1 = nz - 1 if which == 1 else nz
            assert np.array_equal(got[0, 1:j1 + 1, 1:i1 + 1], -got[1, 1:j1 + 1, 1:i1 + 1])
            assert np.array_equal(got[np_ + 1, 1:j1 + 1, 1:i1 + 1], -got[np_, 1:j1 + 1, 1:i1 + 1])
        ora[psi] = got
        ora[psip] = pp
    run.close()


def test_local_nu_matches_the_pinned_formula():
    """vvs_local_nu equals the local Nusselt numbers of the 3D oracle
    (v3_local_nu, the same formula of sect360v2mg.f:318-325, pinned by
    conduction in tests/test_oracle_diagnostics.py) on the same temperature,
    planes k = 0..np, bit for bit"""
    import oracle3 as O3
    nr, nz, np_ = 12, 10, 16
    print("[gpu] sector local Nu", flush=True)
    run = _gpu(nr, nz, np_)
    th = sector_state(nr, nz, np_)["th"]
    th = th + 1e-3 * np.random.default_rng(4).standard_normal(th.shape)
    run.put("th", th)
    a, b = run.local_nu()
    ora = O3.Run3(nr, nz, np_ + 1, **PHYS)
    ora["th"] = th
    oa, ob = ora.local_nu()
    assert np.array_equal(a, oa[:np_ + 1]) and np.array_equal(b, ob[:np_ + 1])
    run.close()


def test_trajectory_200_steps_matches_oracle():
    """200 steps at 8 x 6 x 8 from the reference's start (with the oracle's
    dense solves each step): heat flux and fields within 1e-9."""
    nr, nz, np_ = 8, 6, 8
    print("[gpu] sector 200-step trajectory", flush=True)
    run = _gpu(nr, nz, np_)
    ora = _oracle_run(nr, nz, np_, PHYS)
    for k, v in sector_state(nr, nz, np_).items():
        run.put(k, v)
    for s in range(200):
        tg = run.step(1)
        to = ora.step()
        assert abs(tg - to) <= 1e-9 * abs(to), (s, tg, to)
    for n in ("A", "B", "C", "th", "uel", "vel", "wel"):
        got, want = run.get(n), ora[n]
        scale = max(np.max(np.abs(want)), 1e-300)
        assert np.max(np.abs(got - want)) <= 1e-9 * scale, n
    run.close()
