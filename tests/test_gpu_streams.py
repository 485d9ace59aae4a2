"""The 3D steps run part of their work on a second stream (the heat flux
and v's right-hand side and solve beside w's and u's; in the tank also the
x derivatives beside the y and z ones).  Every kernel's result depends only
on its inputs, so a step must equal, bit for bit, the same routines called
one after another through the routine-level entry points, which use one
stream: a missing wait (a race) would show here as a difference."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES3 = ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "uel", "vel", "wel", "uelp", "velp", "welp")


def _seed(runs, shape, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    for n in NAMES3:
        a = scale * rng.standard_normal(shape)
        for r in runs:
            r.put(n, a)


def _compare(a, b, names):
    for n in names:
        assert np.array_equal(a.get(n), b.get(n)), n


@pytest.mark.parametrize("shape", [(24, 16, 32), (20, 12, 24)])
def test_tank_step_equals_routine_sequence(shape):
    from codesthesis_dir_amd.tank import Tank, TankConfig
    nr, nz, np_ = shape
    print(f"[gpu] tank streams {shape}", flush=True)
    cfg = TankConfig(nr=nr, nz=nz, np=np_, dt=2e-4, Re=37.0, Ptl=5.0, Gr=11.0, lx=1.0, lz=0.8, ly=2.5)
    a, b = Tank(cfg), Tank(cfg)
    _seed((a, b), a.shape, 7 + np_, 0.1)
    for _ in range(3):
        ta = a.step(1)
        b.derivatives()
        b.vorticity()
        b.temperature()
        tb = b.heat()
        for w in (1, 2, 3):
            b.rhs(w)
            b.poisson(w)
        b.update()
        assert ta == tb
    _compare(a, b, NAMES3)
    a.close()
    b.close()


@pytest.mark.parametrize("shape", [(16, 12, 32), (12, 10, 20)])
def test_annulus3d_step_equals_routine_sequence(shape):
    from codesthesis_dir_amd.velvor3 import Annulus3D, Annulus3DConfig
    nr, nz, np_ = shape
    print(f"[gpu] rot3d streams {shape}", flush=True)
    cfg = Annulus3DConfig(nr=nr, nz=nz, np=np_, dt=1e-5, Ra=1e5, Re=100.0, Ptl=7.0, Gr=3.0)
    a, b = Annulus3D(cfg), Annulus3D(cfg)
    _seed((a, b), (np_ + 1, nz + 2, nr + 2), 11 + np_, 0.1)
    for _ in range(3):
        ta = a.step(1)
        b.vorticity()
        b.temperature()
        tb = b.heat()
        b.rhsw()
        b.poisson1()
        b.rhsv()
        b.poisson2()
        b.rhsu()
        b.poisson3()
        b.update()
        assert ta == tb
    _compare(a, b, NAMES3)
    a.close()
    b.close()


@pytest.mark.parametrize("shape", [(12, 10, 16), (10, 8, 12)])
def test_sector_step_equals_routine_sequence(shape):
    from codesthesis_dir_amd.sector import Sector, SectorConfig
    nr, nz, np_ = shape
    print(f"[gpu] sector streams {shape}", flush=True)
    cfg = SectorConfig(nr=nr, nz=nz, np=np_, dt=1e-5, Ra=1e5, Re=100.0, Ptl=7.0, Gr=3.0)
    a, b = Sector(cfg), Sector(cfg)
    _seed((a, b), a.shape, 13 + np_, 0.1)
    for _ in range(3):
        ta = a.step(1)
        b.vorticity(0)
        b.temperature()
        tb = b.heat()
        for w in (1, 2, 3):
            b.rhs(w)
            b.poisson(w)
        b.update()
        assert ta == tb
    _compare(a, b, NAMES3)
    a.close()
    b.close()
