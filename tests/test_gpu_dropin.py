"""The Fortran drop-in (codesthesis_dir_amd/fortran/vv_dropin.F90): a
Fortran program that calls the reference routine names (tests/fortran/
vv_driver.f90, the time loop of velvor2dperiodicrbconv.f:341-388) runs
on the GPU through ISO_C_BINDING and the C ABI, and matches the oracle."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "fortran", "_build", "vv_driver")


@pytest.mark.gpu
def test_fortran_dropin_time_loop():
    assert os.path.exists(DRIVER), "build the drop-in first (__graft_entry__.build())"
    nr, np_, nts, dt, Ra, Pr, width = 32, 64, 60, 1e-5, 1e5, 1.0, 1.0
    run = O.Run(nr, np_, dt, Ra, Pr, width=width)
    run.seed()
    th0 = np.ascontiguousarray(run["th"][:, : nr + 2])
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "th0.bin"), os.path.join(d, "out.bin")
        th0.tofile(fin)
        deck = f"{nr} {np_} {nts} {dt} {Ra} {Pr} {width}\n{fin}\n{fout}\n"
        print(f"[gpu] fortran drop-in driver nr={nr} np={np_} steps={nts}", flush=True)
        res = subprocess.run([DRIVER], input=deck, text=True, capture_output=True, timeout=240)
        assert res.returncode == 0, res.stdout + res.stderr
        raw = np.fromfile(fout, dtype=np.float64)
    n = (nr + 2) * (np_ + 1)
    th, C, uel, vel = (raw[j * n:(j + 1) * n].reshape(np_ + 1, nr + 2) for j in range(4))
    trans = raw[4 * n:].reshape(nts, 2)
    tr_o = []
    for _ in range(nts):
        st, tr, ncl = run.step(mode=1)
        tr_o.append(tr.copy())
    for name, g, ni in (("th", th, nr + 2), ("C", C, nr + 1), ("uel", uel, nr + 1), ("vel", vel, nr + 2)):
        o = run[name][:, :ni]
        err = np.linalg.norm(g[:, :ni] - o) / np.linalg.norm(o)
        print(f"drop-in {name}: rel L2 {err:.3e}")
        assert err <= 1e-9, (name, err)
    assert np.max(np.abs(trans - np.array(tr_o)) / np.abs(np.array(tr_o))) < 1e-10


DRIVER3 = os.path.join(ROOT, "tests", "fortran", "_build", "vv3_driver")


@pytest.mark.gpu
def test_fortran_dropin3_time_loop():
    """The 3D drop-in (vv_dropin3.F90): a Fortran program with rot3dmg.f's
    COMMON blocks (kr = kz = 33, kp = 160) replays the reference's time loop
    (rot3dmg.f:241-253) through the routine names; the fields match the
    oracle's routines with exact solves."""
    import oracle3 as O3
    from test_gpu_3d import oracle_step
    assert os.path.exists(DRIVER3), "build the drop-in first (__graft_entry__.build())"
    nr, nz, np_, nts = 12, 10, 16, 4
    p = dict(dt=1e-4, Ra=1e5, Re=50.0, Ptl=1.0, Gr=3.0)
    ora = O3.Run3(nr, nz, np_, **p)
    rng = np.random.default_rng(8)
    th0 = np.broadcast_to(np.linspace(0.0, 1.0, nr + 2), ora.shape).copy()
    th0 += 0.05 * rng.standard_normal(ora.shape)
    th0[:, :, 0], th0[:, :, -1] = 0.0, 1.0
    th0[0] = th0[-1]
    ora["th"], ora["thp"] = th0, th0
    with tempfile.TemporaryDirectory() as d:
        fin, fout = os.path.join(d, "th0.bin"), os.path.join(d, "out.bin")
        np.ascontiguousarray(th0).tofile(fin)
        deck = f"{nr} {nz} {np_} {nts} {p['dt']} {p['Ra']} {p['Re']} {p['Ptl']} {p['Gr']}\n{fin}\n{fout}\n"
        print(f"[gpu] fortran 3D drop-in driver nr={nr} nz={nz} np={np_} steps={nts}", flush=True)
        res = subprocess.run([DRIVER3], input=deck, text=True, capture_output=True, timeout=240)
        assert res.returncode == 0, res.stdout + res.stderr
        raw = np.fromfile(fout, dtype=np.float64)
    n = int(np.prod(ora.shape))
    th, uel, vel, wel = (raw[j * n:(j + 1) * n].reshape(ora.shape) for j in range(4))
    trans = raw[4 * n:]
    tr_o = [oracle_step(ora) for _ in range(nts)]
    for name, g in (("th", th), ("uel", uel), ("vel", vel), ("wel", wel)):
        o = ora[name]
        err = np.max(np.abs(g - o)) / np.max(np.abs(o))
        print(f"3D drop-in {name}: rel max {err:.3e}")
        assert err <= 1e-10, (name, err)
    assert np.max(np.abs(trans - np.array(tr_o)) / np.abs(np.array(tr_o))) < 1e-10
