"""Parity of the HIP library (through its C ABI) with the oracle and with
the reference fixtures.

Explicit-step routines are compared bit for bit (the kernels keep the
reference's operation order).  Poisson solves are compared with the exact
discrete solution of the reference's residual operator.  Trajectories are
compared in relative L2 after 1000 steps.
"""
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _box(nr, np_, width, **kw):
    from codesthesis_dir_amd import BoxConfig, PeriodicBox
    print(f"[gpu] box nr={nr} np={np_} width={width} {kw}", flush=True)
    return PeriodicBox(BoxConfig(nr=nr, np=np_, width=width, **kw), device=0)


def _annulus(nr, np_, **kw):
    from codesthesis_dir_amd import AnnulusConfig, PeriodicBox
    print(f"[gpu] annulus nr={nr} np={np_} {kw}", flush=True)
    return PeriodicBox(AnnulusConfig(nr=nr, np=np_, **kw), device=0)


def _pair(geom, nr, np_, dt=1e-5, Ra=1e5, Re=100.0, **kw):
    """An oracle run and a GPU run with the same configuration."""
    if geom == 0:
        return O.Run(nr, np_, dt, Ra, 1.0, width=3.0), _box(nr, np_, 3.0, dt=dt, Ra=Ra, **kw)
    return (O.Run(nr, np_, dt, Ra, 1.0, geom=1, r0=0.5, Re=Re),
            _annulus(nr, np_, dt=dt, Ra=Ra, Re=Re, **kw))


def _random_state(run, seed):
    rng = np.random.default_rng(seed)
    nr, np_ = run.nr, run.np
    k = np.arange(np_ + 1)[:, None]
    i = np.arange(run.ld)[None, :]
    smooth = lambda a, b: np.sin(2 * np.pi * a * k / np_ + b * i)
    st = {
        "C": 50 * smooth(2, 0.3) + rng.standard_normal(run["C"].shape),
        "Cp": 50 * smooth(2, 0.31) + rng.standard_normal(run["C"].shape),
        "th": 0.5 + 0.3 * smooth(1, 0.2) + 0.01 * rng.standard_normal(run["C"].shape),
        "thp": 0.5 + 0.3 * smooth(1, 0.21) + 0.01 * rng.standard_normal(run["C"].shape),
        "uel": smooth(3, 0.1) + 0.1 * rng.standard_normal(run["C"].shape),
        "vel": smooth(1, 0.15) + 0.1 * rng.standard_normal(run["C"].shape),
    }
    for n, a in st.items():
        run[n][...] = a
    return st


@pytest.mark.parametrize("geom", [0, 1])
@pytest.mark.parametrize("nr,np_,first", [(16, 32, True), (16, 32, False), (64, 128, False), (1100, 32, False),
                                           (40, 64, False), (96, 336, False), (20, 45, True),
                                           (24, 1024, False), (12, 2048, True), (10, 4096, False),
                                           (16, 12288, False), (10, 16376, True)])
def test_explicit_step_bitexact(nr, np_, first, geom):
    import ctypes
    run, box = _pair(geom, nr, np_)
    st = _random_state(run, seed=nr + np_)
    for n, a in st.items():
        box.put(n, a)
    # oracle
    O.lib().vo_cvorticity(ctypes.byref(run.p), ctypes.byref(run.f), int(first))
    O.lib().vo_temperature(ctypes.byref(run.p), ctypes.byref(run.f), int(first))
    # HIP
    box.cvorticity(first)
    Cn, Cf = box.get("Cn"), box.get("C")
    box.temperature(first)
    thn, thf = box.get("thn"), box.get("th")
    assert np.array_equal(Cn[:, 1:nr], run["Cn"][:, 1:nr])
    assert np.array_equal(Cf, run["C"])
    assert np.array_equal(thn[:, 1:nr + 1], run["thn"][:, 1:nr + 1])
    assert np.array_equal(thf, run["th"])
    # heat on the filtered temperature
    t1, t2 = ctypes.c_double(), ctypes.c_double()
    O.lib().vo_heat(ctypes.byref(run.p), O.P(run["th"]), ctypes.byref(t1), ctypes.byref(t2))
    g1, g2 = box.heat()
    assert (g1, g2) == (t1.value, t2.value)
    # Poisson right-hand sides
    box.rhs()
    vrh, urh = np.zeros_like(run["C"]), np.zeros_like(run["C"])
    O.lib().vo_rhsv(ctypes.byref(run.p), O.P(run["Cn"]), O.P(vrh))
    O.lib().vo_rhsu(ctypes.byref(run.p), O.P(run["Cn"]), O.P(urh))
    assert np.array_equal(box.get("vrh")[:, 1:nr + 1], vrh[:, 1:nr + 1])
    assert np.array_equal(box.get("urh")[:, 1:nr], urh[:, 1:nr])
    # update with a fresh velocity
    rng = np.random.default_rng(1)
    vel = rng.standard_normal(run["C"].shape)
    run["vel"][...] = vel
    box.put("vel", vel)
    ncl_o = O.lib().vo_update(ctypes.byref(run.p), ctypes.byref(run.f))
    ncl_g = box.update()
    assert ncl_g == ncl_o
    for name in ("C", "Cp", "th", "thp"):
        assert np.array_equal(box.get(name)[:, : nr + 2 if name.startswith("th") else nr + 1],
                              run[name][:, : nr + 2 if name.startswith("th") else nr + 1]), name
    box.close()


@pytest.mark.parametrize("nr,np_,width", [(16, 32, 3.0), (64, 128, 1.0), (64, 128, 3.0),
                                          (128, 256, 1.0), (48, 64, 2.0), (64, 128, 0.0),
                                          (256, 256, 0.0), (96, 336, 3.0), (32, 100, 0.0),
                                          (16, 45, 1.0)])
@pytest.mark.parametrize("which", [2, 3])
@pytest.mark.parametrize("direct_cols", [8, 128])
def test_poisson_matches_exact_solution(nr, np_, width, which, direct_cols):
    """width 0 marks the annulus geometry (dp = 2 pi/np, r0 = 0.5).
    direct_cols 8: multigrid down to 8 columns or to the first odd column
    count (336 -> 21, 100 -> 25, 45 at once); 128: a direct level of at
    most 128 columns (two levels at np = 256 and 336, a direct solve at
    np <= 128).  96 x 336 is the size of the reference's own input deck
    (input.s; the box width is 3, rb.f:132-133); 100 and 45 columns are neither powers of two nor multiples
    of the dense transform's 16-column tiles."""
    geom = 1 if width == 0.0 else 0
    if geom:
        box = _annulus(nr, np_, rtol=1e-12, max_cycles=60, Re=100.0, direct_cols=direct_cols)
        dr, dp = 0.5 / nr, 2.0 * np.arccos(-1.0) / np_
    else:
        box = _box(nr, np_, width, rtol=1e-12, max_cycles=60, direct_cols=direct_cols)
        dr, dp = 0.5 / nr, width / np_
    rng = np.random.default_rng(nr * 7 + which)
    ld = nr + 2
    k = np.arange(np_ + 1)[:, None]
    i = np.arange(ld)[None, :]
    rhs = np.ascontiguousarray(np.cos(2 * np.pi * 3 * k / np_) * np.sin(0.4 * i)
                               + 0.2 * rng.standard_normal((np_ + 1, ld)))
    rhs[0] = rhs[np_]
    ni = nr + 2 if which == 2 else nr + 1
    psi = np.zeros((np_ + 1, ld))
    psi[:, 1:ni - 1] = 1e-3 * rng.standard_normal((np_ + 1, ni - 2))
    psi[0] = psi[np_]
    psip = 0.5 * psi
    psi0 = psi.copy()
    st, cyc, rel = box.poisson_host(which, psi, psip, rhs)
    assert st == 0 and rel <= 1e-12 and cyc < 40, (st, cyc, rel)
    ex = np.zeros_like(psi)
    O.poisson_exact(which, ex, np.zeros_like(ex), rhs, nr, np_, dr, dp, geom=geom, r0=0.5)
    err = np.max(np.abs(psi[:, :ni] - ex[:, :ni])) / np.max(np.abs(ex))
    assert err < 1e-10, err
    assert np.array_equal(psip[:, :ni], psi0[:, :ni])       # psip <- old psi
    assert np.all(psi[:, 0] == 0) and np.all(psi[:, ni - 1] == 0)
    assert np.array_equal(psi[0], psi[np_])
    m, _ = O.residual(which, psi, rhs, nr, np_, dr, dp, geom=geom, r0=0.5)
    assert m <= 2e-12 * np.max(np.abs(rhs[1:, 1:ni - 1]))
    box.close()


@pytest.mark.parametrize("nr,np_,width", [(2048, 64, 1.0), (1536, 64, 0.0)])
@pytest.mark.parametrize("which", [2, 3])
@pytest.mark.parametrize("direct_cols", [8, 128])
def test_poisson_long_lines(nr, np_, width, which, direct_cols):
    """Lines longer than 1024 rows (two rows per thread in the line solves).
    The r-operator's coefficients are (nr/0.5)^2 ~ 1e7 here, and evaluating
    the residual in double precision has a relative floor near
    eps (nr/pi)^2 ~ 1e-10, so the solves run at rtol 1e-9 and the solution
    is compared with the exact one at 1e-6 (an indexing error in the long
    lines gives O(1))."""
    geom = 1 if width == 0.0 else 0
    if geom:
        box = _annulus(nr, np_, rtol=1e-9, max_cycles=60, Re=100.0, direct_cols=direct_cols)
        dr, dp = 0.5 / nr, 2.0 * np.arccos(-1.0) / np_
    else:
        box = _box(nr, np_, width, rtol=1e-9, max_cycles=60, direct_cols=direct_cols)
        dr, dp = 0.5 / nr, width / np_
    rng = np.random.default_rng(nr + which)
    ld = nr + 2
    k = np.arange(np_ + 1)[:, None]
    i = np.arange(ld)[None, :]
    rhs = np.ascontiguousarray(np.cos(2 * np.pi * 3 * k / np_) * np.sin(0.05 * i)
                               + 0.2 * rng.standard_normal((np_ + 1, ld)))
    rhs[0] = rhs[np_]
    ni = nr + 2 if which == 2 else nr + 1
    psi, psip = np.zeros((np_ + 1, ld)), np.zeros((np_ + 1, ld))
    st, cyc, rel = box.poisson_host(which, psi, psip, rhs)
    assert st == 0 and rel <= 1e-9, (st, cyc, rel)
    ex = np.zeros_like(psi)
    O.poisson_exact(which, ex, np.zeros_like(ex), rhs, nr, np_, dr, dp, geom=geom, r0=0.5)
    err = np.max(np.abs(psi[:, :ni] - ex[:, :ni])) / np.max(np.abs(ex))
    print(f"nr={nr} which={which} dc={direct_cols}: cycles {cyc} relres {rel:.2e} err {err:.2e}")
    assert err < 1e-6, err
    box.close()


@pytest.mark.parametrize("direct_cols", [1, 8, 4096])
@pytest.mark.parametrize("geom", [0, 1])
def test_poisson_coarse_solve_variants(direct_cols, geom):
    """The exact coarse solve (real DFT over the coarsest level's columns,
    one line solve per mode) at 1 column, at 8 columns, and on the fine
    level itself (direct_cols >= np: the whole solve is direct)."""
    nr, np_ = 96, 128
    if geom:
        box = _annulus(nr, np_, rtol=1e-12, max_cycles=60, Re=100.0, direct_cols=direct_cols)
        dr, dp = 0.5 / nr, 2.0 * np.arccos(-1.0) / np_
    else:
        box = _box(nr, np_, 2.0, rtol=1e-12, max_cycles=60, direct_cols=direct_cols)
        dr, dp = 0.5 / nr, 2.0 / np_
    rng = np.random.default_rng(11)
    rhs = np.ascontiguousarray(rng.standard_normal((np_ + 1, nr + 2)))
    rhs[0] = rhs[np_]
    for which in (2, 3):
        ni = nr + 2 if which == 2 else nr + 1
        psi, psip = np.zeros_like(rhs), np.zeros_like(rhs)
        st, cyc, rel = box.poisson_host(which, psi, psip, rhs)
        assert st == 0 and rel <= 1e-12, (st, cyc, rel)
        if direct_cols >= np_:     # direct solve plus iterative refinement
            assert cyc <= 3, cyc
        ex = np.zeros_like(psi)
        O.poisson_exact(which, ex, np.zeros_like(ex), rhs, nr, np_, dr, dp, geom=geom, r0=0.5)
        err = np.max(np.abs(psi[:, :ni] - ex[:, :ni])) / np.max(np.abs(ex))
        assert err < 1e-10, (which, err)
    box.close()


@pytest.mark.parametrize("geom", [0, 1])
def test_poisson_direct_fft_1024(geom):
    """The default direct solve of the whole plane at 1024 columns (the
    direct level's transforms as FFTs over the columns, one workgroup per
    pair of lines): the exact solve and, at rtol 1e-12, its refinement
    cycles (projection of the residual, added synthesis) against the
    oracle's exact solution."""
    nr, np_ = 24, 1024
    print(f"[gpu] direct solve, FFT transforms, np=1024, geom={geom}", flush=True)
    if geom:
        box = _annulus(nr, np_, rtol=1e-12, max_cycles=60, Re=100.0)
        dr, dp = 0.5 / nr, 2.0 * np.arccos(-1.0) / np_
    else:
        box = _box(nr, np_, 2.0, rtol=1e-12, max_cycles=60)
        dr, dp = 0.5 / nr, 2.0 / np_
    rng = np.random.default_rng(12)
    rhs = np.ascontiguousarray(rng.standard_normal((np_ + 1, nr + 2)))
    rhs[0] = rhs[np_]
    for which in (2, 3):
        ni = nr + 2 if which == 2 else nr + 1
        psi, psip = np.zeros_like(rhs), np.zeros_like(rhs)
        st, cyc, rel = box.poisson_host(which, psi, psip, rhs)
        assert st == 0 and rel <= 1e-12 and cyc <= 3, (st, cyc, rel)
        ex = np.zeros_like(psi)
        O.poisson_exact(which, ex, np.zeros_like(ex), rhs, nr, np_, dr, dp, geom=geom, r0=0.5)
        err = np.max(np.abs(psi[:, :ni] - ex[:, :ni])) / np.max(np.abs(ex))
        assert err < 1e-10, (which, err)
    box.close()


@pytest.mark.parametrize("geom,nr,np_", [(0, 96, 336), (1, 96, 336), (0, 30, 45), (1, 40, 100),
                                         (1, 240, 1792), (0, 24, 3072), (1, 16, 6144), (0, 20, 1200)])
def test_poisson_default_direct_any_np(geom, nr, np_):
    """The default configuration at column counts that are not powers of
    two: one direct solve of the whole plane, against the oracle's exact
    solution, and at rtol 1e-12 its refinement cycles.  Even counts use the
    mixed-radix FFT (336 = 21 * 16, 100 = 25 * 4, 1792 = 7 * 256 as in the
    annulus deck input1e81e5pr1-1.s, 1200 = 75 * 16; 3072 and 6144 take
    the 32-register combination), the odd 45 the dense transforms (no
    Nyquist line)."""
    if geom:
        box = _annulus(nr, np_, rtol=1e-12, max_cycles=60, Re=100.0)
        dr, dp = 0.5 / nr, 2.0 * np.arccos(-1.0) / np_
    else:
        box = _box(nr, np_, 3.0, rtol=1e-12, max_cycles=60)
        dr, dp = 0.5 / nr, 3.0 / np_
    assert box.nlev == 1
    rng = np.random.default_rng(np_)
    k = np.arange(np_ + 1)[:, None]
    rhs = np.ascontiguousarray(np.cos(2 * np.pi * (np_ // 2) * k / np_) * np.ones((1, nr + 2))
                               + rng.standard_normal((np_ + 1, nr + 2)))
    rhs[0] = rhs[np_]
    for which in (2, 3):
        ni = nr + 2 if which == 2 else nr + 1
        psi, psip = np.zeros_like(rhs), np.zeros_like(rhs)
        st, cyc, rel = box.poisson_host(which, psi, psip, rhs)
        assert st == 0 and rel <= 1e-12 and cyc <= 3, (st, cyc, rel)
        ex = np.zeros_like(psi)
        O.poisson_exact(which, ex, np.zeros_like(ex), rhs, nr, np_, dr, dp, geom=geom, r0=0.5)
        err = np.max(np.abs(psi[:, :ni] - ex[:, :ni])) / np.max(np.abs(ex))
        assert err < 1e-10, (which, err)
        assert np.array_equal(psi[0], psi[np_])
    box.close()


@pytest.mark.parametrize("nr,np_", [(16, 12288), (12, 16376)])
def test_trajectory_above_8192_columns(nr, np_):
    """Above 8192 columns the compact tiles no longer fit in LDS (their
    tables move to global memory) and the default Poisson path is the
    V-cycle with a 128-column direct level: 5 steps against the oracle with
    exact solves, relative L2 of every field <= 1e-9."""
    width, dt, Ra, Pr = 3.0, 1e-6, 1e5, 1.0
    run = O.Run(nr, np_, dt, Ra, Pr, width=width)
    run.seed()
    box = _box(nr, np_, width, dt=dt, Ra=Ra, Pr=Pr, rtol=1e-12, max_cycles=60)
    assert box.nlev > 1
    box.put("th", run["th"])
    box.start()
    info = box.step(5)
    assert info.status == 0
    for _ in range(5):
        st, tr, ncl = run.step(mode=1)
        assert st == 0
    for name in ("th", "C", "uel", "vel"):
        g, o = box.get(name), run[name]
        err = np.linalg.norm(g - o) / np.linalg.norm(o)
        assert err <= 1e-9, (name, err)
    box.close()


def test_trajectory_reference_deck_size():
    """The reference's own input deck (input.s: nr = 96, np = 336,
    Ra = 1e7, Pr = 0.71, dt = 1e-7; the box width is 3, rb.f:132-133),
    seeded temperature, 200 steps with the default direct solves against the oracle with exact
    solves: relative L2 of every field <= 1e-9, wall fluxes to 1e-10."""
    nr, np_, width, dt, Ra, Pr = 96, 336, 3.0, 1e-7, 1e7, 0.71
    run = O.Run(nr, np_, dt, Ra, Pr, width=width)
    run.seed()
    box = _box(nr, np_, width, dt=dt, Ra=Ra, Pr=Pr, rtol=1e-10, max_cycles=60)
    box.put("th", run["th"])
    box.start()
    info = box.step(200)
    assert info.status == 0
    trans = []
    for _ in range(200):
        st, tr, ncl = run.step(mode=1)
        assert st == 0
        trans.append(tr.copy())
    for name in ("th", "C", "uel", "vel", "Cp", "thp"):
        g, o = box.get(name), run[name]
        err = np.linalg.norm(g - o) / np.linalg.norm(o)
        print(f"{name}: rel L2 {err:.3e}")
        assert err <= 1e-9, (name, err)
    hist = box.heat_history(200)
    trans = np.array(trans)
    assert np.max(np.abs(hist - trans) / np.abs(trans)) < 1e-10
    box.close()


def test_unsteady_high_ra_time_averaged_nusselt():
    """Unsteady convection at Ra = 4e6 (box of nr = 24, np = 72, width 1.5,
    so square cells; Pr = 0.71, dt = 2e-6), 30000 steps from the seeded
    start; the wall Nusselt numbers oscillate (about 4.7-6.1 after the
    transient).  Their time averages over steps 10000-30000 at both walls
    agree with the oracle's (exact solves) within 0.5 %, the criterion for
    the high-Ra runs.  At this resolution a copy of the oracle started 1e-12
    away stays within 1e-8 of it, i.e. the run is time-dependent, not
    chaotic: the series also agree step by step, which is printed."""
    nr, np_, width, dt, Ra, Pr, N = 24, 72, 1.5, 2e-6, 4e6, 0.71, 30000
    run = O.Run(nr, np_, dt, Ra, Pr, width=width)
    run.seed()
    box = _box(nr, np_, width, dt=dt, Ra=Ra, Pr=Pr, rtol=1e-10, max_cycles=60)
    box.put("th", run["th"])
    box.start()
    info = box.step(N)
    assert info.status == 0
    hist = box.heat_history(N)
    trans = np.zeros((N, 2))
    for n in range(N):
        st, tr, ncl = run.step(mode=1)
        assert st == 0
        trans[n] = tr
    nu_g = np.abs(hist) * (0.5 / width)        # rb.f:384-388
    nu_o = np.abs(trans) * (0.5 / width)
    assert np.all(np.isfinite(nu_g))
    mg, mo = nu_g[N // 3:].mean(axis=0), nu_o[N // 3:].mean(axis=0)
    rel = np.abs(mg / mo - 1)
    step_rel = np.max(np.abs(nu_g - nu_o) / np.abs(nu_o))
    print(f"time-averaged Nu GPU {mg}, oracle {mo}, rel {rel}; largest step-by-step rel {step_rel:.2e}; "
          f"std over the window {nu_o[N // 3:].std(axis=0)}")
    assert np.all(rel < 5e-3), rel
    assert nu_o[N // 3:, 0].std() > 0.05        # the flow is unsteady
    box.close()


@pytest.mark.parametrize("direct_cols", [16, 128])
def test_trajectory_1000_steps_box_129x65(direct_cols):
    """Periodic RB box 129x65 nodes (nr=64, np=128), 2:1 (width 1, height
    0.5), Ra=1e5, Pr=1, dt=1e-5, seeded temperature, 1000 steps with the
    Poisson solves at 1e-10 relative residual.  Relative L2 of every field
    against the oracle with exact solves must be <= 1e-9.  direct_cols 16:
    V-cycles over 128/64/32 columns and a direct 16-column level; 128: the
    default, a direct solve."""
    nr, np_, width = 64, 128, 1.0
    run = O.Run(nr, np_, 1e-5, 1e5, 1.0, width=width)
    run.seed()
    box = _box(nr, np_, width, rtol=1e-10, max_cycles=60, direct_cols=direct_cols)
    box.put("th", run["th"])
    box.start()
    info = box.step(1000)
    assert info.status == 0
    trans = []
    for _ in range(1000):
        st, tr, ncl = run.step(mode=1)
        assert st == 0
        trans.append(tr.copy())
    for name in ("th", "C", "uel", "vel", "Cp", "thp"):
        g, o = box.get(name), run[name]
        err = np.linalg.norm(g - o) / np.linalg.norm(o)
        print(f"{name}: rel L2 {err:.3e}")
        assert err <= 1e-9, (name, err)
    hist = box.heat_history(1000)
    trans = np.array(trans)
    assert np.max(np.abs(hist - trans) / np.abs(trans)) < 1e-10
    box.close()


@pytest.mark.parametrize("name,snap", [("traj_16x32.npz", 20), ("traj_64x128.npz", 1000)])
def test_trajectory_against_reference_fixture(golden_dir, name, snap):
    """Against the reference's own run (flang build, absolute Poisson stop
    2e-6): temperature and wall fluxes agree to round-off; the velocity
    and vorticity differ by the reference's solver tolerance (about 1e-5
    relative, measured with the exact-solve oracle)."""
    g = np.load(os.path.join(golden_dir, name))
    nr, np_ = int(g["nr"]), int(g["np"])
    box = _box(nr, np_, float(g["width"]), rtol=1e-12, max_cycles=60)
    box.put("th", g["th0"])
    box.start()
    box.step(snap)
    for f, tol in (("th", 1e-10), ("C", 5e-5), ("uel", 5e-5), ("vel", 5e-5)):
        ref = g[f"s{snap}_{f}"]
        a = box.get(f)[:, : ref.shape[1]]
        err = np.linalg.norm(a - ref) / np.linalg.norm(ref)
        print(f"{name} {f}: {err:.3e}")
        assert err < tol, (f, err)
    hist = box.heat_history(snap)
    assert np.max(np.abs(hist - g["trans"]) / np.abs(g["trans"])) < 1e-10
    box.close()


@pytest.mark.parametrize("direct_cols", [16, 128])
def test_annulus_trajectory_300_steps(direct_cols):
    """Rotating-annulus plane (velvor2dnewsch3.4.f), 65x129 nodes, Ra=1e7,
    Re=sqrt(1e5), dt=1e-6, 300 steps: relative L2 <= 1e-9 against the oracle
    with exact solves.  The solves run at 1e-12: at 1e-10 the azimuthal
    velocity differs by 3.8e-9 (measured with V-cycles), the solver
    tolerance times the conditioning of the annulus operator."""
    nr, np_ = 64, 128
    Re = float(np.sqrt(1e5))
    run, box = _pair(1, nr, np_, dt=1e-6, Ra=1e7, Re=Re, rtol=1e-12, max_cycles=60,
                     direct_cols=direct_cols)
    run.seed()
    box.put("th", run["th"])
    box.start()
    info = box.step(300)
    assert info.status == 0
    for _ in range(300):
        st, tr, ncl = run.step(mode=1)
        assert st == 0
    for name in ("th", "C", "uel", "vel", "Cp", "thp"):
        g, o = box.get(name), run[name]
        err = np.linalg.norm(g - o) / np.linalg.norm(o)
        print(f"annulus {name}: rel L2 {err:.3e}")
        assert err <= 1e-9, (name, err)
    box.close()


@pytest.mark.parametrize("name,snap", [("ann_16x32.npz", 20), ("ann_64x128.npz", 300)])
def test_annulus_against_reference_fixture(golden_dir, name, snap):
    """Against the annulus reference's own run (absolute Poisson stop 1e-5,
    ann.f:2093): temperature and wall fluxes to round-off, the velocity and
    vorticity to the reference's solver tolerance."""
    g = np.load(os.path.join(golden_dir, name))
    nr, np_ = int(g["nr"]), int(g["np"])
    box = _annulus(nr, np_, dt=float(g["dt"]), Ra=float(g["Ra"]), Re=float(g["Re"]), rtol=1e-12,
                   max_cycles=60)
    box.put("th", g["th0"])
    box.start()
    box.step(snap)
    for f, tol in (("th", 1e-9), ("C", 1e-5), ("uel", 1e-5), ("vel", 1e-5)):
        ref = g[f"s{snap}_{f}"]
        a = box.get(f)[:, : ref.shape[1]]
        err = np.linalg.norm(a - ref) / np.linalg.norm(ref)
        print(f"{name} {f}: {err:.3e}")
        assert err < tol, (f, err)
    hist = box.heat_history(snap)
    assert np.max(np.abs(hist - g["trans"]) / np.abs(g["trans"])) < 1e-9
    box.close()


def test_concurrent_calls_on_one_run():
    """poisson2 and poisson3 on one run from two threads at once, as the
    reference's main program calls them from two OpenMP sections
    (rb.f:370-380): the run's lock serialises the calls, so every result is
    bit-identical to the sequential one.  ctypes releases the interpreter
    lock during the calls, so they do overlap."""
    import threading
    nr, np_ = 64, 128
    box = _box(nr, np_, 3.0, rtol=1e-12, max_cycles=60)
    rng = np.random.default_rng(5)
    rhs = {w: np.ascontiguousarray(rng.standard_normal((np_ + 1, nr + 2))) for w in (2, 3)}
    for w in (2, 3):
        rhs[w][0] = rhs[w][np_]
    psi0 = {w: np.ascontiguousarray(1e-3 * rng.standard_normal((np_ + 1, nr + 2))) for w in (2, 3)}

    def solve(w, out):
        psi, psip = psi0[w].copy(), 0.5 * psi0[w]
        out[w] = (box.poisson_host(w, psi, psip, rhs[w].copy()), psi, psip)

    ref = {}
    for w in (2, 3):
        solve(w, ref)
    for rep in range(20):
        got = {}
        bar = threading.Barrier(2)

        def run(w):
            bar.wait()
            solve(w, got)
        th = [threading.Thread(target=run, args=(w,)) for w in (2, 3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        for w in (2, 3):
            assert got[w][0][0] == 0, (rep, w, got[w][0])
            assert np.array_equal(got[w][1], ref[w][1]), (rep, w)
            assert np.array_equal(got[w][2], ref[w][2]), (rep, w)
    box.close()


@pytest.mark.parametrize("nr,np_", [(16, 20000), (3000, 64), (4, 64), (16, 6), (16, 4097), (16, 16382)])
def test_sizes_the_kernels_cannot_hold_are_rejected(nr, np_):
    """vv_create refuses sizes whose lines do not fit in LDS (np above 16383,
    nr above 2048), sizes below 8, and a direct level without an FFT of
    more than 4096 columns (2 c^2 dense weights): an odd np above 4095, or
    16382 = 2 x 8191, whose halving stops at 8191 columns."""
    import codesthesis_dir_amd._lib as L
    with pytest.raises(L.VVError) as e:
        _box(nr, np_, 3.0)
    assert e.value.code == L.VV_EINVAL


def test_nonfinite_rhs_is_an_error():
    from codesthesis_dir_amd import _lib as L
    nr, np_ = 16, 32
    box = _box(nr, np_, 1.0)
    rhs = np.zeros((np_ + 1, nr + 2))
    rhs[5, 7] = np.nan
    psi, psip = np.zeros_like(rhs), np.zeros_like(rhs)
    with pytest.raises(L.VVError) as e:
        box.poisson_host(2, psi, psip, rhs)
    assert e.value.code == L.VV_ENONFINITE
    box.close()


def test_cycle_cap_reports_not_converged():
    nr, np_ = 32, 64
    box = _box(nr, np_, 1.0, rtol=1e-15, max_cycles=2, direct_cols=4)
    rng = np.random.default_rng(0)
    rhs = np.ascontiguousarray(rng.standard_normal((np_ + 1, nr + 2)))
    psi, psip = np.zeros_like(rhs), np.zeros_like(rhs)
    st, cyc, rel = box.poisson_host(3, psi, psip, rhs)
    assert st == 1 and cyc == 2 and rel > 1e-15
    box.close()


def test_odd_sweep_count_is_rejected():
    """The two-colour sweeps alternate between two buffers per level, so
    nu1 + nu2 must be even."""
    from codesthesis_dir_amd import _lib as L
    with pytest.raises(L.VVError) as e:
        _box(16, 32, 1.0, nu1=1, nu2=2)
    assert e.value.code == L.VV_EINVAL


def test_deferred_solves_refine_a_cycle_that_misses_rtol():
    """A deferred step whose first direct cycle misses rtol is refined on
    the device by the queued second cycle: with rtol between the one-cycle
    residuals of the steps, a run of deferred steps equals, bit for bit,
    the run whose every solve is synchronised (and runs as many cycles as
    it needs), and reports no failure."""
    from codesthesis_dir_amd.ic import box_temperature
    print("[gpu] deferred refinement", flush=True)
    nr, np_ = 32, 64
    th0 = box_temperature(nr, np_, 0.5, 1.0, 1e-3, 3)
    probe = _box(nr, np_, 1.0, rtol=1e-10)
    probe.put("th", th0)
    probe.start()
    probe.profile(True)
    rel = []
    for _ in range(12):
        info = probe.step(1)
        rel += [info.relres_v, info.relres_u]
    probe.close()
    rel = sorted(rel)
    rtol = (rel[len(rel) // 4] * rel[3 * len(rel) // 4]) ** 0.5   # some steps need a second cycle
    a = _box(nr, np_, 1.0, rtol=rtol, max_cycles=2)
    b = _box(nr, np_, 1.0, rtol=rtol, max_cycles=2)
    for box in (a, b):
        box.put("th", th0)
        box.start()
    b.profile(True)                 # every solve synchronised
    cycles, status = [], []
    for _ in range(12):
        info = b.step(1)
        cycles += [info.cycles_v, info.cycles_u]
        status.append(info.status)
    assert max(cycles) == 2 and max(status) == 0
    a.step(1)                       # synchronised; the rest deferred once a solve takes one cycle
    info = a.step(11)
    assert info.status == 0
    for f in ("C", "th", "uel", "vel"):
        assert np.array_equal(a.get(f), b.get(f)), f
    a.close()
    b.close()


def test_deferred_solves_report_a_nonfinite_rhs():
    """After a first synchronised solve converged in one cycle, vv_step
    queues its direct solves without waiting (a device check records
    failures, read once per call): the fields equal those of a run whose
    every solve is synchronised (per-kernel timing on), and a NaN that
    reaches a later right-hand side is still reported as VV_ENONFINITE."""
    from codesthesis_dir_amd import _lib as L
    from codesthesis_dir_amd.ic import box_temperature
    print("[gpu] deferred solves", flush=True)
    nr, np_ = 32, 64
    a = _box(nr, np_, 1.0)
    b = _box(nr, np_, 1.0)
    th0 = box_temperature(nr, np_, 0.5, 1.0, 1e-3, 3)
    for box in (a, b):
        box.put("th", th0)
        box.start()
    a.step(6)                       # one synchronised solve, five deferred
    b.profile(True)                 # timed kernels: every solve synchronised
    b.step(6)
    b.profile(False)
    for f in ("C", "th", "uel", "vel"):
        assert np.array_equal(a.get(f), b.get(f)), f
    th = a.get("th")
    th[10, 5] = np.nan
    a.put("th", th)
    with pytest.raises(L.VVError) as e:
        a.step(3)
    assert e.value.code == L.VV_ENONFINITE
    a.close()
    b.close()
