"""The oracle against the reference's own outputs (tests/golden/*.npz,
written by tests/golden/make_golden.py from the flang build of
velvor2dperiodicrbconv.f).  The oracle keeps the reference's evaluation
order, so equality is bit-for-bit."""
import os

import numpy as np
import pytest

import oracle as O


def _load(golden_dir, name):
    path = os.path.join(golden_dir, name)
    if not os.path.exists(path):
        pytest.skip(f"fixture {name} missing")
    return np.load(path)


def test_banded_solvers_bitexact(golden_dir):
    g = _load(golden_dir, "banded.npz")
    for n in (7, 33, 129):
        y = O.tdma(g[f"tdma_{n}_a"], g[f"tdma_{n}_b"], g[f"tdma_{n}_c"], g[f"tdma_{n}_r"])
        assert np.array_equal(y, g[f"tdma_{n}_y"])
        y = O.ptdma(g[f"ptdma_{n}_a"], g[f"ptdma_{n}_b"], g[f"ptdma_{n}_c"], g[f"ptdma_{n}_r"])
        assert np.array_equal(y, g[f"ptdma_{n}_y"])
        y = O.pdma(*(g[f"pdma_{n}_{k}"] for k in "abcder"))
        assert np.array_equal(y, g[f"pdma_{n}_y"])


def test_poisson_literal_bitexact(golden_dir):
    g = _load(golden_dir, "poisson_16x32.npz")
    nr, np_ = int(g["nr"]), int(g["np"])
    run = O.Run(nr, np_, 1e-5, 1e5, 1.0, width=3.0)
    import ctypes
    for which in (2, 3):
        ni = g[f"p{which}_rhs"].shape[1]
        pad = lambda a: np.ascontiguousarray(np.pad(a, ((0, 0), (0, run.ld - ni))))
        psi, psip, rhs = pad(g[f"p{which}_psi_in"]), pad(g[f"p{which}_psip_in"]), pad(g[f"p{which}_rhs"])
        it = ctypes.c_int(0)
        st = O.lib().vo_poisson_ref(ctypes.byref(run.mg), which, O.P(psi), O.P(psip), O.P(rhs),
                                    nr, np_, ctypes.c_double(2e-6), 500, ctypes.byref(it))
        assert st == 0
        assert np.array_equal(psi[:, :ni], g[f"p{which}_psi"])
        assert np.array_equal(psip[:, :ni], g[f"p{which}_psip"])


@pytest.mark.parametrize("name,snaps", [("traj_16x32.npz", (1, 2, 20)),
                                        ("traj_64x128.npz", (1, 1000))])
def test_trajectory_bitexact(golden_dir, name, snaps):
    g = _load(golden_dir, name)
    nr, np_ = int(g["nr"]), int(g["np"])
    run = O.Run(nr, np_, float(g["dt"]), float(g["Ra"]), float(g["Pr"]), width=float(g["width"]))
    run.seed(r0=float(g["r0"]), amp=float(g["amp"]), seed=int(g["seed"]))
    assert np.array_equal(run["th"], g["th0"])
    trans = g["trans"]
    for s in range(1, max(snaps) + 1):
        st, tr, ncl = run.step(mode=0, tolcon=2e-6)
        assert st == 0 and ncl == 0
        assert np.array_equal(tr, trans[s - 1]), s
        if s in snaps:
            for name_f in ("th", "thp", "C", "Cp", "uel", "vel"):
                ref = g[f"s{s}_{name_f}"]
                assert np.array_equal(run[name_f][:, : ref.shape[1]], ref), (s, name_f)


@pytest.mark.parametrize("name,snaps", [("ann_16x32.npz", (1, 20)), ("ann_64x128.npz", (1, 300))])
def test_annulus_trajectory_bitexact(golden_dir, name, snaps):
    """The rotating-annulus plane (velvor2dnewsch3.4.f) restated by the
    oracle against the reference's own run."""
    g = _load(golden_dir, name)
    nr, np_ = int(g["nr"]), int(g["np"])
    run = O.Run(nr, np_, float(g["dt"]), float(g["Ra"]), float(g["Pr"]), geom=1,
                r0=float(g["r0"]), Re=float(g["Re"]))
    run.seed(r0=float(g["r0"]), amp=float(g["amp"]), seed=int(g["seed"]))
    assert np.array_equal(run["th"], g["th0"])
    trans = g["trans"]
    for s in range(1, max(snaps) + 1):
        st, tr, ncl = run.step(mode=0)
        assert st == 0 and ncl == 0
        assert np.array_equal(tr, trans[s - 1]), s
        if s in snaps:
            for name_f in ("th", "thp", "C", "Cp", "uel", "vel"):
                ref = g[f"s{s}_{name_f}"]
                assert np.array_equal(run[name_f][:, : ref.shape[1]], ref), (s, name_f)
