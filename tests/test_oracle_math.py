"""Pins of the oracle that do not depend on the reference build: each check
is fixed by mathematics (a dense solve, a closed form, an exactness
property), chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails."""
import ctypes

import numpy as np
import pytest

import oracle as O


# ------------------------------------------------------------ banded solvers

def test_tdma_is_the_tridiagonal_solve():
    rng = np.random.default_rng(0)
    for n in (3, 10, 65):
        a, c = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        b, r = rng.uniform(3, 5, n), rng.uniform(-1, 1, n)
        M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        assert np.allclose(O.tdma(a, b, c, r), np.linalg.solve(M, r), rtol=1e-13, atol=1e-14)


def test_ptdma_is_the_cyclic_solve():
    """rb.f:3066 couples row 1 to row n through a(1) and row n to row 1
    through c(n)."""
    rng = np.random.default_rng(1)
    for n in (5, 17, 129):
        a, c = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        b, r = rng.uniform(3, 5, n), rng.uniform(-1, 1, n)
        M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        M[0, n - 1] = a[0]
        M[n - 1, 0] = c[n - 1]
        assert np.allclose(O.ptdma(a, b, c, r), np.linalg.solve(M, r), rtol=1e-12, atol=1e-13)


def test_pdma_is_the_pentadiagonal_solve():
    rng = np.random.default_rng(2)
    for n in (4, 12, 64):
        a, b = rng.uniform(-0.3, 0.3, n), rng.uniform(-1, 1, n)
        c = rng.uniform(4, 6, n)
        d, e = rng.uniform(-1, 1, n), rng.uniform(-0.3, 0.3, n)
        M = np.diag(c) + np.diag(b[1:], -1) + np.diag(a[2:], -2) + np.diag(d[:-1], 1) + np.diag(e[:-2], 2)
        r = rng.uniform(-1, 1, n)
        assert np.allclose(O.pdma(a, b, c, d, e, r), np.linalg.solve(M, r), rtol=1e-12, atol=1e-13)


# ------------------------------------------------------- compact derivatives

def _run(nr=16, np_=32, width=3.0):
    return O.Run(nr, np_, 1e-5, 1e5, 1.0, width=width)


def test_compact_phi_derivative_is_the_np_plus_1_cyclic_system():
    """Ch (rb.f:833-851): 1-4-1 cyclic system of size np+1 on data of period
    np, rhs 3(C_{k+1}-C_{k-1})/dp with k=0 and k=np both using C_1, C_{np-1}.
    Rebuilt here as a dense (np+1)x(np+1) solve."""
    run = _run()
    nr, np_, dp = run.nr, run.np, run.p.dp
    rng = np.random.default_rng(3)
    C = rng.standard_normal((np_ + 1, run.ld))
    C[np_] = C[0]
    run["C"][...] = C
    run["Cp"][...] = C
    O.lib().vo_cvorticity(ctypes.byref(run.p), ctypes.byref(run.f), 1)
    n = np_ + 1
    M = 4 * np.eye(n) + np.eye(n, k=1) + np.eye(n, k=-1)
    M[0, n - 1] = 1.0
    M[n - 1, 0] = 1.0
    for i in (1, 5, nr - 1):
        col = C[:, i]
        rhs = np.empty(n)
        for k in range(n):
            kp = 1 if k in (0, np_) else k + 1
            km = np_ - 1 if k in (0, np_) else k - 1
            rhs[k] = 3.0 * (col[kp] - col[km]) / dp
        assert np.allclose(run["B"][:, i], np.linalg.solve(M, rhs), rtol=1e-12, atol=1e-12)


def test_compact_phi_derivative_seam_error():
    """The (np+1)-cyclic system is not the period-np circulant: on a smooth
    periodic field the derivative is 4th-order accurate away from the seam
    and visibly wrong next to it."""
    np_ = 64
    run = O.Run(8, np_, 1e-5, 1e5, 1.0, width=1.0)
    k = np.arange(np_ + 1)
    x = k * run.p.dp
    f = np.sin(2 * np.pi * x)
    run["C"][...] = f[:, None]
    run["Cp"][...] = run["C"]
    O.lib().vo_cvorticity(ctypes.byref(run.p), ctypes.byref(run.f), 1)
    d = run["B"][:, 3]
    exact = 2 * np.pi * np.cos(2 * np.pi * x)
    mid = slice(np_ // 4, 3 * np_ // 4)
    assert np.max(np.abs(d[mid] - exact[mid])) < 1e-5
    assert np.max(np.abs(d - exact)) > 1e-4


def test_half_derivative_of_staggered_temperature():
    """thr/thh at vorticity points (rb.f:854-897) carry a factor 1.5 on a
    1-4-1 system: half of the derivative of the averaged temperature."""
    nr, np_ = 32, 64
    run = O.Run(nr, np_, 1e-5, 1e5, 1.0, width=1.0)
    dp = run.p.dp
    k = np.arange(np_ + 1)
    th = np.sin(2 * np.pi * k * dp)[:, None] * np.ones((1, run.ld))
    run["th"][...] = th
    run["thp"][...] = th
    O.lib().vo_cvorticity(ctypes.byref(run.p), ctypes.byref(run.f), 1)
    thh = run["Bn"][:, nr // 2]
    xh = (k + 0.5) * dp
    exact = 2 * np.pi * np.cos(2 * np.pi * xh)
    mid = slice(np_ // 4, 3 * np_ // 4)
    ratio = thh[mid] / exact[mid]
    assert np.allclose(ratio, 0.5, atol=2e-3)


# ------------------------------------------------------------------- heat

def test_heat_weights_are_fornberg_one_sided():
    """The wall flux (rb.f:1483-1495) uses the one-sided first derivative
    at the wall through points at 0, h/2, 3h/2, 5h/2, 7h/2 with weights
    -352/105, 35/8, -35/24, 21/40, -5/56 (truncated to 8-9 digits)."""
    x = np.array([0.0, 0.5, 1.5, 2.5, 3.5])
    V = np.vander(x, 5, increasing=True).T          # rows: powers
    rhs = np.zeros(5)
    rhs[1] = 1.0
    w = np.linalg.solve(V, rhs)
    assert np.allclose(w, [-352 / 105, 35 / 8, -35 / 24, 21 / 40, -5 / 56], rtol=1e-12)
    nr, np_ = 16, 32
    run = O.Run(nr, np_, 1e-5, 1e5, 1.0, width=2.0)
    dr = run.p.dr
    r = np.empty(run.ld)
    r[0] = 0.0
    r[1:nr + 1] = (np.arange(1, nr + 1) - 0.5) * dr
    r[nr + 1] = nr * dr
    a, b, c = 0.3, 1.7, -2.1
    prof = a + b * r + c * r ** 2 + 0.4 * r ** 3
    th = np.tile(prof, (np_ + 1, 1))
    t1, t2 = ctypes.c_double(), ctypes.c_double()
    O.lib().vo_heat(ctypes.byref(run.p), O.P(th), ctypes.byref(t1), ctypes.byref(t2))
    H = nr * dr
    width = np_ * run.p.dp
    # the truncated decimals leave a weight-sum defect of about 1e-8,
    # which multiplies the wall value by 1/dr
    tol = 3e-8 * np.max(np.abs(prof)) / dr * width
    assert abs(t1.value - b * width) < tol
    slope_top = b + 2 * c * H + 1.2 * H ** 2
    # trans2 is the one-sided derivative taken inward from the top wall
    assert abs(t2.value - (-slope_top) * width) < tol


# ------------------------------------------------------------------ update

def test_wall_vorticity_is_exact_for_quadratics():
    """Cn(0) = (3 v1 - v2/3)/dr (rb.f:1588) is the wall derivative of a
    profile vanishing at the wall, exact for quadratics."""
    nr, np_ = 16, 8
    run = O.Run(nr, np_, 1e-5, 1e5, 1.0, width=1.0)
    dr = run.p.dr
    H = nr * dr
    x = (np.arange(run.ld) - 0.5) * dr
    a, b = 1.3, -4.0
    v_bot = a * x + b * x ** 2
    xt = H - x
    v_top = a * xt + b * xt ** 2
    vel = np.zeros((np_ + 1, run.ld))
    vel[:, 1:3] = v_bot[1:3]
    vel[:, nr - 1:nr + 1] = v_top[nr - 1:nr + 1]
    run["vel"][...] = vel
    O.lib().vo_update(ctypes.byref(run.p), ctypes.byref(run.f))
    assert np.allclose(run["C"][:, 0], a, rtol=1e-12)
    assert np.allclose(run["C"][:, nr], -a, rtol=1e-12)


# ---------------------------------------------------------------- Poisson

def _dense_operator(which, nr, np_, dr, dp, ld):
    """Columns of L from the oracle's residual applied to unit vectors."""
    ihi = nr if which == 2 else nr - 1
    idx = [(i, k) for k in range(1, np_ + 1) for i in range(1, ihi + 1)]
    n = len(idx)
    M = np.zeros((n, n))
    zero = np.zeros((np_ + 1, ld))
    for j, (i, k) in enumerate(idx):
        psi = np.zeros((np_ + 1, ld))
        psi[k, i] = 1.0
        if k == np_:
            psi[0, i] = 1.0
        _, out = O.residual(which, psi, zero, nr, np_, dr, dp)
        M[:, j] = [-out[kk, ii] for (ii, kk) in idx]
    return M, idx


@pytest.mark.parametrize("which", [2, 3])
def test_residual_operator_matches_stencil(which):
    """The residual operator is the 5-point Laplacian with the relax2
    closure rows (-2, 5/3, -1/5)/dr^2 at i=1 and i=nr (rb.f:2326-2333)."""
    nr, np_, dr, dp = 6, 8, 0.1, 0.2
    M, idx = _dense_operator(which, nr, np_, dr, dp, nr + 2)
    pos = {ik: j for j, ik in enumerate(idx)}
    i, k = 1, 3
    row = M[pos[(i, k)]]
    if which == 2:
        assert np.isclose(row[pos[(1, 3)]], -2 / dr**2 - 2 / dp**2)
        assert np.isclose(row[pos[(2, 3)]], 5 / 3 / dr**2)
        assert np.isclose(row[pos[(3, 3)]], -0.2 / dr**2)
        row = M[pos[(nr, 5)]]
        assert np.isclose(row[pos[(nr - 1, 5)]], 5 / 3 / dr**2)
        assert np.isclose(row[pos[(nr - 2, 5)]], -0.2 / dr**2)
    else:
        assert np.isclose(row[pos[(1, 3)]], -2 / dr**2 - 2 / dp**2)
        assert np.isclose(row[pos[(2, 3)]], 1 / dr**2)
    # periodic wrap in phi
    row = M[pos[(2, 1)]]
    assert np.isclose(row[pos[(2, np_)]], 1 / dp**2)
    assert np.isclose(row[pos[(2, 2)]], 1 / dp**2)


@pytest.mark.parametrize("which", [2, 3])
def test_exact_poisson_against_dense_solve(which):
    nr, np_ = 8, 12
    dr, dp = 0.5 / nr, 3.0 / np_
    ld = nr + 2
    M, idx = _dense_operator(which, nr, np_, dr, dp, ld)
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal((np_ + 1, ld))
    rhs[0] = rhs[np_]
    f = np.array([rhs[k, i] for (i, k) in idx])
    x = np.linalg.solve(M, f)
    psi = np.zeros((np_ + 1, ld))
    psip = np.zeros_like(psi)
    O.poisson_exact(which, psi, psip, rhs, nr, np_, dr, dp)
    got = np.array([psi[k, i] for (i, k) in idx])
    assert np.allclose(got, x, rtol=1e-11, atol=1e-13 * np.max(np.abs(x)))
    assert np.array_equal(psi[0], psi[np_])


@pytest.mark.parametrize("geom", [0, 1])
@pytest.mark.parametrize("which", [2, 3])
def test_literal_solver_converges_to_exact(which, geom):
    """Tightening the reference's absolute stop drives its FMG/line-GS
    iteration (box: rb.f; annulus: velvor2dnewsch3.4.f) to the exact
    discrete solution of the separable solver."""
    nr, np_ = 32, 64
    run = O.Run(nr, np_, 1e-5, 1e5, 1.0, width=3.0, geom=geom, r0=0.5, Re=100.0)
    rng = np.random.default_rng(9)
    k = np.arange(np_ + 1)[:, None]
    i = np.arange(run.ld)[None, :]
    rhs = np.ascontiguousarray(np.cos(2 * np.pi * k / np_) * np.sin(0.2 * i) + 0.05 * rng.standard_normal((np_ + 1, run.ld)))
    rhs[np_] = rhs[0]
    psi_e = np.zeros_like(rhs)
    O.poisson_exact(which, psi_e, np.zeros_like(rhs), rhs, nr, np_, run.p.dr, run.p.dp,
                    geom=geom, r0=0.5)
    psi = np.zeros_like(rhs)
    psip = np.zeros_like(rhs)
    it = ctypes.c_int()
    tol = 1e-11 * np.max(np.abs(rhs))
    st = O.lib().vo_poisson_ref(ctypes.byref(run.mg), which, O.P(psi), O.P(psip), O.P(rhs),
                                nr, np_, ctypes.c_double(tol), 2000, ctypes.byref(it))
    assert st == 0
    ni = nr + 2 if which == 2 else nr + 1
    err = np.max(np.abs(psi[:, :ni] - psi_e[:, :ni])) / np.max(np.abs(psi_e))
    assert err < 1e-9
    m, _ = O.residual(which, psi_e, rhs, nr, np_, run.p.dr, run.p.dp, geom=geom, r0=0.5)
    assert m < 1e-11 * np.max(np.abs(rhs))


def test_seed_stream_matches_python_splitmix():
    nr, np_ = 4, 6
    run = O.Run(nr, np_, 1e-5, 1e5, 1.0, width=1.0)
    run.seed(r0=0.5, amp=1e-3, seed=123)
    u = O.splitmix64_uniform(123, nr * np_).reshape(np_, nr)
    r = 0.5 + (np.arange(1, nr + 1) - 0.5) * run.p.dr
    expect = 2 * (r - 0.5)[None, :] + 1e-3 * (2 * u - 1)
    assert np.array_equal(run["th"][:np_, 1:nr + 1], expect)
    assert np.array_equal(run["th"][np_], run["th"][0])
    assert np.all(run["th"][:, 0] == 0.0) and np.all(run["th"][:, nr + 1] == 1.0)
