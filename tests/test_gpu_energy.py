"""Kinetic-energy integrals on the GPU (vv_energy, vv3_energy; no
counterpart in the reference, the definition is in include/vvmi.h and
include/vvmi3.h) against the oracle's sums, whose ranges and weights the
CPU tests pin by closed forms.  The GPU adds in another order, so the bar
is 1e-13 relative; a repeated call is bit-for-bit the same."""
import numpy as np
import pytest

import oracle as O
import oracle3 as O3

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("geom,nr,np_", [(0, 64, 128), (1, 48, 128), (1, 8, 2048), (0, 300, 16)])
def test_energy_2d_matches_oracle(geom, nr, np_):
    from codesthesis_dir_amd import BoxConfig, PeriodicBox
    from codesthesis_dir_amd.velvor import AnnulusConfig
    print(f"[gpu] energy 2d geom={geom} nr={nr} np={np_}", flush=True)
    kw = dict(nr=nr, np=np_, dt=1e-5, Ra=1e4, Pr=1.0)
    cfg = AnnulusConfig(Re=10.0, **kw) if geom else BoxConfig(width=1.0, **kw)
    box = PeriodicBox(cfg)
    ora = O.Run(nr, np_, 1e-5, 1e4, 1.0, width=1.0, geom=geom, r0=0.5, Re=10.0)
    rng = np.random.default_rng(nr + np_)
    for n in ("uel", "vel"):
        a = rng.standard_normal((np_ + 1, nr + 2))
        box.put(n, a)
        ora[n][...] = a
    ke = box.energy()
    want = ora.energy()
    assert abs(ke - want) <= 1e-13 * want, (ke, want)
    assert box.energy() == ke
    box.close()


@pytest.mark.parametrize("shape", [(16, 12, 24), (9, 7, 14), (40, 30, 1100)])
def test_energy_3d_matches_oracle(shape):
    from codesthesis_dir_amd.velvor3 import Annulus3D, Annulus3DConfig
    nr, nz, np_ = shape
    print(f"[gpu] energy 3d {shape}", flush=True)
    run = Annulus3D(Annulus3DConfig(nr=nr, nz=nz, np=np_, dt=1e-4, Ra=1e4, Re=30.0, Ptl=1.0, Gr=0.0))
    ora = O3.Run3(nr, nz, np_, 1e-4, 1e4, 30.0, 1.0, 0.0)
    rng = np.random.default_rng(np_)
    for n in ("uel", "vel", "wel"):
        a = rng.standard_normal(ora.shape)
        run.put(n, a)
        ora[n] = a
    ke = run.energy()
    want = ora.energy()
    assert abs(ke - want) <= 1e-13 * want, (ke, want)
    assert run.energy() == ke
    run.close()


@pytest.mark.parametrize("shape", [(16, 12, 24), (33, 20, 160)])
def test_local_nu_and_modes_match_oracle(shape):
    """vv3_local_nu equals the oracle bit for bit (same operations per
    plane); vv3_modes within 1e-12 of the oracle's sums, relative to the
    sum of the plane sums' magnitudes (the GPU adds in another order)."""
    from codesthesis_dir_amd.velvor3 import Annulus3D, Annulus3DConfig
    nr, nz, np_ = shape
    print(f"[gpu] diagnostics 3d {shape}", flush=True)
    run = Annulus3D(Annulus3DConfig(nr=nr, nz=nz, np=np_, dt=1e-4, Ra=1e4, Re=30.0, Ptl=1.0, Gr=0.0))
    ora = O3.Run3(nr, nz, np_, 1e-4, 1e4, 30.0, 1.0, 0.0)
    th = np.random.default_rng(np_ + 1).standard_normal(ora.shape)
    run.put("th", th)
    ora["th"] = th
    a, b = run.local_nu()
    oa, ob = ora.local_nu()
    assert np.array_equal(a, oa) and np.array_equal(b, ob)
    c, s = run.modes(8)
    oc, os_ = ora.modes(8)
    scale = run.cfg.zmax / nz * 0.5 / nr * 2 * np.pi / np_ * np.sum(np.abs(th))
    assert np.max(np.abs(c - oc)) <= 1e-12 * scale and np.max(np.abs(s - os_)) <= 1e-12 * scale
    run.close()
