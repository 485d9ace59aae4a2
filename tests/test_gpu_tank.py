"""The 3D horizontal-convection tank on the GPU (include/vvtank.h) against
the oracle (oracle/vvt_oracle.c, pinned in tests/test_oracle_tank.py):

* the compact derivatives, the explicit routines, the heat fluxes, the
  right-hand sides and the update evaluate the oracle's expressions in its
  order: bit for bit;
* the solves: within 1e-11 of the oracle's dense solve, with its output
  conventions (walls, ghost planes, psip) exactly;
* whole steps: within 1e-10 of the oracle's;
* the 512 x 256 x 256 tank (np = 512 along the tank, Ra = 1e9, Pr = 5, Ta = 1e8, uniform T
  plus seeded noise): after 20 steps every solve satisfies its operator to
  rounding, the temperature stays within the wall values, the fields stay
  finite and the ghost planes mirror.
"""
import numpy as np
import pytest

import oracle_tank as OT

pytestmark = pytest.mark.gpu

PHYS = dict(dt=2e-4, Re=37.0, Ptl=5.0, Gr=11.0)
NAMES = ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "uel", "vel", "wel", "An", "Bn", "Cn", "thn",
         "uelp", "velp", "welp", "urh", "vrh", "wrh")


def _pair(nr, nz, np_, seed, scale=1.0, **kw):
    from codesthesis_dir_amd.tank import Tank, TankConfig
    p = dict(PHYS)
    p.update(kw)
    cfg = TankConfig(nr=nr, nz=nz, np=np_, lx=1.0, lz=0.8, ly=2.5, **p)
    run = Tank(cfg)
    ora = OT.Tank(nr, nz, np_, dt=cfg.dt, Re=cfg.Re, Ptl=cfg.Ptl, Gr=cfg.Gr, lx=1.0, lz=0.8, ly=2.5)
    rng = np.random.default_rng(seed)
    for n in NAMES:
        a = scale * rng.standard_normal(ora.shape)
        ora[n] = a
        run.put(n, a)
    return run, ora


@pytest.mark.parametrize("shape", [(8, 6, 10), (13, 9, 7), (24, 16, 40)])
def test_explicit_routines_bitexact(shape):
    nr, nz, np_ = shape
    print(f"[gpu] tank explicit {shape}", flush=True)
    run, ora = _pair(nr, nz, np_, seed=nr + 3 * nz + 7 * np_)
    run.derivatives()
    d = ora.derivatives()
    for q in range(12):
        assert np.array_equal(run.derivative(q), d[q]), q
    run.vorticity()
    run.temperature()
    for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
        ora.call(rt)
    for n in ("An", "Bn", "Cn", "thn"):
        assert np.array_equal(run.get(n), ora[n]), n
    assert run.heat() == ora.heat()
    for which, rhs in ((1, "wrh"), (2, "vrh"), (3, "urh")):
        run.rhs(which)
        ora.call({1: "rhsw", 2: "rhsv", 3: "rhsu"}[which])
        assert np.array_equal(run.get(rhs), ora[rhs]), rhs
    run.update()
    ora.call("update")
    for n in ("A", "Ap", "An", "B", "Bp", "Bn", "C", "Cp", "Cn", "th", "thp", "thn"):
        assert np.array_equal(run.get(n), ora[n]), n
    run.close()


@pytest.mark.parametrize("shape", [(8, 6, 10), (13, 9, 40), (24, 16, 17)])
def test_step_explicit_part_bitexact(shape):
    """A whole step runs the four updates in one pass: after one step the
    new A, B, C and th at their updated points (the solves and the update
    touch only velocities, walls and ghosts) and the heat fluxes are the
    oracle's, bit for bit."""
    nr, nz, np_ = shape
    print(f"[gpu] tank step explicit {shape}", flush=True)
    run, ora = _pair(nr, nz, np_, seed=11 + nr + 5 * nz + np_)
    tg = run.step(1)
    to = ora.step()
    assert tg == to
    inner = {"A": np.s_[1:np_, 1:nz, 1:nr + 1], "B": np.s_[1:np_ + 1, 1:nz, 1:nr],
             "C": np.s_[1:np_, 1:nz + 1, 1:nr], "th": np.s_[1:np_ + 1, 1:nz + 1, 1:nr + 1]}
    for n, sl in inner.items():
        assert np.array_equal(run.get(n)[sl], ora[n][sl]), n
    run.close()


@pytest.mark.parametrize("shape", [(6, 6, 2100), (40, 70, 6)])
def test_derivatives_long_lines_bitexact(shape):
    """Lines longer than the tables staged in LDS (y lines of 2100 cells) and
    x / z lines spanning several checkpoint segments, bit for bit."""
    nr, nz, np_ = shape
    print(f"[gpu] tank derivatives {shape}", flush=True)
    run, ora = _pair(nr, nz, np_, seed=5 + nr + nz + np_)
    run.derivatives()
    d = ora.derivatives()
    for q in range(12):
        assert np.array_equal(run.derivative(q), d[q]), q
    run.close()


@pytest.mark.parametrize("which", [1, 2, 3])
@pytest.mark.parametrize("shape", [(8, 6, 10), (7, 9, 6), (35, 6, 8), (40, 6, 7)])
def test_poisson_matches_dense_solve(which, shape):
    """Including x lines of several 16-row segments (35 and 40 cells), with
    y modes by FFT (np = 8) and by GEMM."""
    nr, nz, np_ = shape
    print(f"[gpu] tank poisson {which} {shape}", flush=True)
    run, ora = _pair(nr, nz, np_, seed=which)
    psi, psip, rhs = OT.PSI[which]
    old = ora[psi].copy()
    run.poisson(which)
    ora.poisson(which)
    got, want = run.get(psi), ora[psi]
    assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want))
    i1, j1, k1 = ora.ranges(which)
    inner = np.zeros(ora.shape, bool)
    inner[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = True
    ghost = np.zeros(ora.shape, bool)
    if which != 2:
        ghost[[0, np_ + 1], 1:j1 + 1, 1:i1 + 1] = True
    assert np.all(got[~inner & ~ghost] == 0.0)
    if which != 2:
        assert np.array_equal(got[0], got[1]) and np.array_equal(got[np_ + 1], got[np_])
    assert np.array_equal(run.get(psip), old)
    m, _ = ora.residual(which, got, ora[rhs])
    assert m <= 1e-10 * np.max(np.abs(ora[rhs])) * max(nr, nz, np_) ** 2
    run.close()


def _conduction(nr, nz, np_, seed, t0=0.5, amp=1e-4):
    rng = np.random.default_rng(seed)
    th = t0 + amp * (2.0 * rng.uniform(size=(np_ + 2, nz + 2, nr + 2)) - 1.0)
    return th


def test_trajectory_matches_oracle():
    nr, nz, np_ = 10, 8, 12
    print("[gpu] tank trajectory", flush=True)
    from codesthesis_dir_amd.tank import Tank, TankConfig
    p = dict(dt=1e-4, Re=100.0, Ptl=5.0, Gr=2e4)
    cfg = TankConfig(nr=nr, nz=nz, np=np_, lx=1.0, lz=1.0, ly=4.0, **p)
    run = Tank(cfg)
    ora = OT.Tank(nr, nz, np_, lx=1.0, lz=1.0, ly=4.0, **p)
    th = _conduction(nr, nz, np_, 20261206, amp=1e-2)
    for n in ("th", "thp"):
        run.put(n, th)
        ora[n] = th
    for s in range(6):
        tg = run.step(1)
        to = ora.step()
        for a, b in zip(tg, to):
            # the bottom flux is near zero (insulated): bound by the flux scale
            assert abs(a - b) <= 1e-10 * max(abs(b), 1.0), (s, tg, to)
    for n in ("A", "B", "C", "th", "uel", "vel", "wel", "Ap", "thp", "welp"):
        got, want = run.get(n), ora[n]
        scale = max(np.max(np.abs(want)), 1e-300)
        assert np.max(np.abs(got - want)) <= 1e-10 * scale, n
    assert abs(run.energy() - ora.energy()) <= 1e-12 * ora.energy()
    run.close()


def test_tank_512x256x256_grid_properties():
    """The 512 x 256 x 256 tank (np = 512 along the tank, nr = nz = 256),
    Ra = Gr Pr = 1e9 with Pr = 5, Ta = 1e8 (Re = 1e4), T = 0.5 + U(+-1e-4)."""
    from codesthesis_dir_amd.tank import Tank, TankConfig
    nr, nz, np_ = 256, 256, 512
    print("[gpu] tank 512x256x256 grid", flush=True)
    cfg = TankConfig(nr=nr, nz=nz, np=np_, dt=1e-7, Re=1e4, Ptl=5.0, Gr=2e8, lx=1.0, lz=1.0, ly=4.0)
    run = Tank(cfg)
    th = _conduction(nr, nz, np_, 20261206)
    run.put("th", th)
    run.put("thp", th)
    del th
    t1, t2 = run.step(20)
    assert np.isfinite(t1) and np.isfinite(t2)
    ora = OT.Tank(nr, nz, np_, dt=cfg.dt, Re=cfg.Re, Ptl=cfg.Ptl, Gr=cfg.Gr, lx=1.0, lz=1.0, ly=4.0)
    th = run.get("th")
    assert th.min() >= -1e-3 and th.max() <= 1.0 + 1e-3
    for which in (1, 2, 3):
        psi, _, rhs = OT.PSI[which]
        got, f = run.get(psi), run.get(rhs)
        assert np.all(np.isfinite(got))
        m, _ = ora.residual(which, got, f)
        scale = np.max(np.abs(f)) + np.max(np.abs(got)) * 12.0 / min(ora.p.dr, ora.p.dz, ora.p.dp) ** 2
        assert m <= 1e-13 * scale, (which, m, scale)
        if which != 2:
            assert np.array_equal(got[0], got[1]) and np.array_equal(got[np_ + 1], got[np_])
        else:
            assert np.all(got[[0, np_, np_ + 1]] == 0.0)
        del got, f
    e = run.energy()
    assert np.isfinite(e) and e > 0.0
    run.close()
