"""The 3D oracle (oracle/vv3_oracle.c, rot3dmg.f restated) against the
reference: exact equality with the golden fixtures written from the
reference build (tests/golden/make_golden3.py), with the live reference
build where it is present, and the exact Poisson solve against the
reference's residual operators."""
import os

import numpy as np
import pytest

import oracle3 as O3
import refdrv3 as R3

HERE = os.path.dirname(os.path.abspath(__file__))
NR, NZ, NP = 16, 12, 24
PARAMS = dict(dt=1e-4, Ra=1e4, Re=30.0, Ptl=1.0, Gr=5.0)


def _load(name):
    return np.load(os.path.join(HERE, "golden", name))


def _run_from(g, model=0):
    run = O3.Run3(NR, NZ, NP, model=model, **PARAMS)
    for n in O3.FIELDS:
        run[n] = g["in_" + n]
    return run


@pytest.mark.parametrize("fixture,model", [("r3_explicit.npz", 0), ("r3h_explicit.npz", 1)])
def test_explicit_routines_match_reference_fixture(fixture, model):
    """rot3dmg.f (model 0) and rot3dmghor.f (model 1, its gravity terms in
    all three vorticity components, swapped wall temperatures and two-wall
    heat flux): every routine's output equals the reference build's."""
    g = _load(fixture)
    run = _run_from(g, model)
    for rt, name in [("avorticity", "An"), ("bvorticity", "Bn"), ("cvorticity", "Cn"), ("temperature", "thn")]:
        run.call(rt)
        assert np.array_equal(run[name], g[rt]), rt
    if model == 0:
        assert run.heat() == float(g["heat"])
    else:
        assert run.heat2() == tuple(float(x) for x in g["heat"])
    for fn, name in [("rhsw", "wrh"), ("rhsv", "vrh"), ("rhsu", "urh")]:
        run.call(fn)
        assert np.array_equal(run[name], g[name]), fn
    run.call("update")
    for n in ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "An", "Bn", "Cn", "thn"):
        assert np.array_equal(run[n], g["upd_" + n]), n


def test_fixture_exercises_every_hybrid_branch():
    """The fixture's velocities put the cell Peclet number above 2 (the
    hybrid-upwind switch, e.g. rot3dmg.f:397-406) in every direction and
    both signs, so the upwind branches are part of the equality above."""
    g = _load("r3_explicit.npz")
    dr, dz, dp = 0.5 / NR, 0.5 / NZ, 2 * np.pi / NP
    u = 0.5 * (g["in_uel"][:, 1:-1, 1:] + g["in_uel"][:, 1:-1, :-1])
    w = 0.5 * (g["in_wel"][:, 1:, 1:-1] + g["in_wel"][:, :-1, 1:-1])
    v = 0.5 * (g["in_vel"][1:, 1:-1, 1:-1] + g["in_vel"][:-1, 1:-1, 1:-1])
    for vel, h in ((u, dr), (w, dz), (v, dp * 0.5)):
        assert np.sum(vel * h > 2) > 10 and np.sum(vel * h < -2) > 10


def test_residual_operators_match_reference_fixture():
    g = _load("r3_resid.npz")
    run = O3.Run3(NR, NZ, NP, **PARAMS)
    for w in (1, 2, 3):
        i1 = NR if w < 3 else NR - 1
        j1 = NZ - 1 if w == 1 else NZ
        _, res = run.residual(w, g[f"psif{w}"], g[f"rhs{w}"])
        sl = (slice(1, NP + 1), slice(1, j1 + 1), slice(1, i1 + 1))
        assert np.array_equal(res[sl], g[f"rc{w}"][sl]), w


@pytest.mark.parametrize("which", [1, 2, 3])
def test_exact_solve_zeroes_the_reference_residual(which):
    """v3_poisson_exact solves the operator the residual test above pins:
    its residual is at the rounding level."""
    run = O3.Run3(NR, NZ, NP, **PARAMS)
    rng = np.random.default_rng(which)
    rhs = rng.standard_normal(run.shape)
    rhs[0] = rhs[NP]
    psi = run.poisson_exact(which, rhs)
    m, _ = run.residual(which, psi, rhs)
    assert m < 1e-12 * np.max(np.abs(rhs)) * (NR * NR), m
    assert np.array_equal(psi[0], psi[NP])


def test_exact_solve_against_dense_solve():
    """A tiny grid assembled densely from the residual operator and solved by
    numpy: the same solution."""
    nr, nz, np_ = 4, 4, 6
    run = O3.Run3(nr, nz, np_, **PARAMS)
    rng = np.random.default_rng(7)
    for which in (1, 2, 3):
        i1 = nr if which < 3 else nr - 1
        j1 = nz - 1 if which == 1 else nz
        idx = [(k, j, i) for k in range(1, np_ + 1) for j in range(1, j1 + 1) for i in range(1, i1 + 1)]
        n = len(idx)
        A = np.zeros((n, n))
        zero = np.zeros(run.shape)
        for c, (k, j, i) in enumerate(idx):
            e = np.zeros(run.shape)
            e[k, j, i] = 1.0
            if k == np_:
                e[0, j, i] = 1.0
            _, r = run.residual(which, e, zero)
            A[:, c] = [-r[kk, jj, ii] for (kk, jj, ii) in idx]
        rhs = rng.standard_normal(run.shape)
        rhs[0] = rhs[np_]
        x = np.linalg.solve(A, np.array([rhs[k, j, i] for (k, j, i) in idx]))
        psi = run.poisson_exact(which, rhs)
        got = np.array([psi[k, j, i] for (k, j, i) in idx])
        assert np.max(np.abs(got - x)) < 1e-12 * np.max(np.abs(x)), which


@pytest.mark.parametrize("model", [0, 1])
def test_explicit_routines_match_live_reference(model):
    """A second seeded state, checked against the reference build itself
    (librot3d.so or librot3dhor.so)."""
    if not R3.available(model):
        pytest.skip("reference build under oracle/_ref absent")
    run = O3.Run3(NR, NZ, NP, ldi=R3.KR1, ldj=R3.KZ1, model=model, **PARAMS)
    ref = R3.Reference3(model)
    ref.setup(run.p)
    rng = np.random.default_rng(99)
    K = NP + 1
    for n in O3.FIELDS:
        scale = 300.0 if n in ("uel", "vel", "wel", "uelp", "velp", "welp") else 1.0
        a = scale * rng.standard_normal((K, R3.KZ1, R3.KR1))
        a[0] = a[NP]
        run[n] = a
        ref.f[n][...] = 0.0
        ref.f[n][:K] = a
    for rt, name in [("avorticity", "An"), ("bvorticity", "Bn"), ("cvorticity", "Cn"), ("temperature", "thn")]:
        run.call(rt)
        ref.call(rt)
        assert np.array_equal(run[name], ref.f[name][:K]), rt
    assert run.heat2() == ref.heat2()
