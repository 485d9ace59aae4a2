"""The sector oracle (oracle/vvs_oracle.c: sect360v2mg.f restated from its
text, since that file does not compile here) pinned by what does not depend
on retyping it:

* away from the barrier its routines and operators agree with the oracle
  of rot3dmg.f (oracle/vv3_oracle.c, itself equal bit for bit to the
  compiled rot3dmg.f, tests/test_oracle3.py), on the same arrays;
* the barrier rows are exact where the mathematics fixes them: the wall
  closures of B, th and A reproduce second and first derivatives of
  quadratics, the insulated barrier planes of th those of quadratics with
  zero slope at the wall, the barrier vorticity the slope of a linear
  velocity;
* the three Poisson operators: Neumann closures annihilate constants, the
  phi (and for u the z) closures have the closed-form eigenvectors of the
  Neumann (cosines at cell centres) or Dirichlet (sines) second
  difference, the w operator is self-adjoint under the weight r (flux
  form), and the exact solve zeroes each residual.
"""
import numpy as np
import pytest

import oracle3 as O3
import oracle_sector as OS

PHYS = dict(dt=1e-4, Ra=1e4, Re=30.0, Ptl=1.0, Gr=5.0)


def _pair(nr=8, nz=7, np_=10, seed=3, scale=1.0):
    """The sector state and a rot3dmg.f state with np+1 planes on the same
    arrays and the sector's dp: rot3dmg.f's routines then read the same
    neighbours as the sector's wherever they do not wrap."""
    s = OS.Sector(nr, nz, np_, **PHYS)
    a = O3.Run3(nr, nz, np_ + 1, **PHYS)
    a.p.dp = s.p.dp
    rng = np.random.default_rng(seed)
    for n in O3.FIELDS:
        x = scale * rng.standard_normal(s.shape)
        s[n] = x
        a[n] = x
    return s, a


def test_interior_matches_rot3dmg():
    nr, nz, np_ = 8, 7, 10
    s, a = _pair(nr, nz, np_)
    # velocities of order 1 with dr = 1/16: the hybrid switch stays off in
    # some cells and on in others
    for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
        s.call(rt)
        a.call(rt)
    # A: every point but the outer-wall row (its Ar is one-sided in the
    # sector, dum15, and repeats the inner row's formula in rot3dmg.f)
    assert np.array_equal(s["An"][1:np_, 1:nz, 1:nr], a["An"][1:np_, 1:nz, 1:nr])
    assert not np.array_equal(s["An"][1:np_, 1:nz, nr], a["An"][1:np_, 1:nz, nr])
    # B: the planes between the barrier rows
    assert np.array_equal(s["Bn"][2:np_, 1:nz, 1:nr], a["Bn"][2:np_, 1:nz, 1:nr])
    assert np.array_equal(s["Cn"][1:np_, 1:nz + 1, 1:nr], a["Cn"][1:np_, 1:nz + 1, 1:nr])
    # th: the sector divides the central phi difference by 2 dp, rot3dmg.f
    # multiplies by 0.5 and divides by dp: equal to rounding
    got, want = s["thn"][2:np_, :, 1:nr + 1], a["thn"][2:np_, :, 1:nr + 1]
    assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want))


def test_rhs_w_u_interior_match_rot3dmg():
    nr, nz, np_ = 8, 7, 10
    s, a = _pair(nr, nz, np_, seed=5)
    for rt in ("rhsw", "rhsu", "rhsv"):
        s.call(rt)
        a.call(rt)
    assert np.array_equal(s["vrh"][1:np_, 1:nz + 1, 1:nr + 1], a["vrh"][1:np_, 1:nz + 1, 1:nr + 1])
    for n, sl in (("wrh", np.s_[2:np_, 1:nz, 1:nr + 1]), ("urh", np.s_[2:np_, 1:nz + 1, 1:nr])):
        got, want = s[n][sl], a[n][sl]
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want)), n
    # at k = 1 the sector leaves out the barrier plane's An and Cn
    # (sect360v2mg.f:1020-1021, 1090-1091): rot3dmg.f's formula with those zero
    s2, a2 = _pair(nr, nz, np_, seed=5)
    for f in (s2, a2):
        f["An"][0] = 0.0
        f["Cn"][0] = 0.0
        f.call("rhsw")
        f.call("rhsu")
    for n, sl in (("wrh", np.s_[1, 1:nz, 1:nr + 1]), ("urh", np.s_[1, 1:nz + 1, 1:nr])):
        got, want = s2[n][sl], a2[n][sl]
        assert np.max(np.abs(got - want)) <= 1e-13 * np.max(np.abs(want)), n


def _xphi(np_, dp):
    """phi positions of the cell planes k = 0..np+1: the barrier faces at 0
    and np dp, cells at (k - 1/2) dp"""
    x = (np.arange(np_ + 2) - 0.5) * dp
    x[0], x[-1] = 0.0, np_ * dp
    return x


def _cn(nr, nz, np_, p, i, j, k, r, centred):
    """the diagonal halves cnr, cnz, cnp of the DuFort-Frankel stencils: -1/h^2
    (over r^2 in phi) inside, -2.5/h^2 on a row half a spacing from a wall
    (with -+0.5/(dr r) in r), -4/h^2 on an insulated wall plane"""
    dr, dz, dp = p.dr, p.dz, p.dp
    if centred and i == 1:
        cnr = -2.5 / dr ** 2 + 0.5 / dr / r
    elif centred and i == nr:
        cnr = -2.5 / dr ** 2 - 0.5 / dr / r
    else:
        cnr = -1.0 / dr ** 2
    cnz = {1: -2.5, nz: -2.5, 0: -4.0, nz + 1: -4.0}.get(j, -1.0) / dz ** 2
    cnp = {1: -2.5, np_: -2.5, 0: -4.0, np_ + 1: -4.0}.get(k, -1.0) / dp ** 2 / r ** 2
    return cnr, cnz, cnp


def test_b_barrier_closures_exact_on_quadratics():
    """B = q(phi) (constant in r and z), v = V0 uniform, everything else 0:
    Bn = B + (q'' - q - r V0 q' / Ptl) / (r^2 cf1) at every plane k = 1..np,
    the barrier rows included."""
    nr, nz, np_ = 6, 6, 9
    s = OS.Sector(nr, nz, np_, **PHYS)
    p = s.p
    x = _xphi(np_, p.dp)
    c0, c1, c2 = 0.3, -0.7, 0.45
    q, q1, q2 = c0 + c1 * x + c2 * x * x, c1 + 2 * c2 * x, 2 * c2 * np.ones_like(x)
    V0 = 0.05
    s["B"] = q[:, None, None]
    s["Bp"] = q[:, None, None]
    s["vel"] = V0
    s.call("bvorticity")
    for k in range(1, np_ + 1):
        for i in (1, 3, nr - 1):
            r = p.r0 + i * p.dr
            cnh = (-2.5 if k in (1, np_) else -1.0) / p.dp ** 2 / r ** 2
            cf1 = 1.0 / (p.Ptl * p.dt) + 1.0 / (2 * r * r) + 1.0 / p.dr ** 2 + 1.0 / p.dz ** 2 - cnh
            want = q[k] + (q2[k] / r ** 2 - q[k] / r ** 2 - V0 * q1[k] / r / p.Ptl) / cf1
            got = s["Bn"][k, 3, i]
            assert abs(got - want) <= 1e-12 * abs(want), (k, i, got, want)


def test_a_outer_row_closure_exact_on_quadratics():
    """A = q(r) (walls at r0 and rmax, cells between), u = U0 uniform:
    An = A + (q'' + q'/r - q/r^2 - U0 q' / Ptl) / cf1 on every row, the
    one-sided rows i = 1 and i = nr included."""
    nr, nz, np_ = 7, 6, 8
    s = OS.Sector(nr, nz, np_, **PHYS)
    p = s.p
    xr = p.r0 + (np.arange(nr + 2) - 0.5) * p.dr
    xr[0], xr[-1] = p.r0, p.r0 + nr * p.dr
    c0, c1, c2 = 0.2, 0.9, -0.6
    q, q1, q2 = c0 + c1 * xr + c2 * xr * xr, c1 + 2 * c2 * xr, 2 * c2
    U0 = 0.1
    s["A"] = q[None, None, :]
    s["Ap"] = q[None, None, :]
    s["uel"] = U0
    s.call("avorticity")
    for i in range(1, nr + 1):
        r = xr[i]
        cnr, cnz, cnh = _cn(nr, nz, np_, p, i, 3, 4, r, True)
        cf1 = 1.0 / (p.Ptl * p.dt) + 1.0 / (2 * r * r) - cnr - cnz - cnh
        want = q[i] + (q2 + q1[i] / r - q[i] / r ** 2 - U0 * q1[i] / p.Ptl) / cf1
        got = s["An"][4, 3, i]
        assert abs(got - want) <= 1e-12 * abs(want), (i, got, want)


@pytest.mark.parametrize("wall", [0, 1])
def test_temperature_barrier_rows_exact_on_quadratics(wall):
    """th = q(phi), v = V0: thn = th + (q''/r^2 - ULT) / cf1 with
    ULT = V0 q'/r inside and on the row k = 1, 0 on the insulated planes;
    q has zero slope at the barrier face that is checked (k = 0 for wall 0,
    k = np+1 for wall 1).  On the row k = np the reference advects with the
    slope of the row k = np-1 (sect360v2mg.f:787 writes its one-sided phi
    derivative into thz): ULT = V0 q'(phi_{np-1}) / r."""
    nr, nz, np_ = 6, 6, 9
    s = OS.Sector(nr, nz, np_, **PHYS)
    p = s.p
    x = _xphi(np_, p.dp)
    x0 = 0.0 if wall == 0 else np_ * p.dp
    c0, c2 = 0.4, 0.8
    q, q1, q2 = c0 + c2 * (x - x0) ** 2, 2 * c2 * (x - x0), 2 * c2
    V0 = 0.05
    s["th"] = q[:, None, None]
    s["thp"] = q[:, None, None]
    s["vel"] = V0
    s.call("temperature")
    ks = [0, 1, 2, 5, np_ - 1, np_] if wall == 0 else [1, 4, np_ - 1, np_, np_ + 1]
    for k in ks:
        for i in (1, 3, nr):
            for j in (0, 1, 3, nz, nz + 1):
                r = p.r0 + (i - 0.5) * p.dr
                cnr, cnz, cnp = _cn(nr, nz, np_, p, i, j, k, r, True)
                cf1 = 1.0 / p.dt - cnr - cnz - cnp
                moving = 1 <= j <= nz and 1 <= k <= np_
                slope = q1[np_ - 1] if k == np_ else q1[k]
                ult = V0 * slope / r if moving else 0.0
                want = q[k] + (q2 / r ** 2 - ult) / cf1
                got = s["thn"][k, j, i]
                assert abs(got - want) <= 1e-12 * abs(want), (k, j, i, got, want)


def test_barrier_vorticity_is_the_slope_of_a_linear_velocity():
    """w = c (phi - phi_face) and u likewise, zero on the barrier faces:
    An = c / r and Cn = -c / r on both faces, r = r0 + (i - 1/2) dr as the
    reference writes it for both (sect360v2mg.f:967, 972)."""
    nr, nz, np_ = 6, 5, 8
    s = OS.Sector(nr, nz, np_, **PHYS)
    p = s.p
    x = _xphi(np_, p.dp)
    c = 0.37
    for face, k in ((0.0, 0), (np_ * p.dp, np_)):
        s["wel"] = (c * (x - face))[:, None, None]
        s["uel"] = (c * (x - face))[:, None, None]
        s.call("update")
        for i in range(1, nr + 1):
            r = p.r0 + (i - 0.5) * p.dr
            assert np.allclose(s["A"][k, 1:nz + 1, i], c / r, rtol=1e-13, atol=0)
            if i < nr:
                assert np.allclose(s["C"][k, 1:nz + 1, i], -c / r, rtol=1e-13, atol=0)


def _apply(s, which, psi):
    """L psi on the unknowns (rhs 0: the residual is -L psi)"""
    _, res = s.residual(which, psi, np.zeros(s.shape))
    return -res


def test_operators_match_rot3dmg_away_from_changed_rows():
    nr, nz, np_ = 8, 7, 10
    s, a = _pair(nr, nz, np_, seed=9)
    rng = np.random.default_rng(2)
    for which in (1, 2, 3):
        i1, j1, k1 = s.ranges(which)
        psi = np.zeros(s.shape)
        psi[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = rng.standard_normal((k1, j1, i1))
        ls = _apply(s, which, psi)
        _, ra = a.residual(which, psi, np.zeros(a.shape))
        la = -ra
        if which == 1:   # w: the sector's r rows are Neumann, rot3dmg.f's extrapolate to 0
            sl = np.s_[2:k1, 1:j1 + 1, 2:i1]
        elif which == 2:   # v: the same operator, Dirichlet barrier faces
            sl = np.s_[1:k1 + 1, 1:j1 + 1, 1:i1 + 1]
        else:   # u: the sector's z rows are Neumann
            sl = np.s_[2:k1, 2:j1, 1:i1 + 1]
        # (v: rot3dmg.f with np+1 planes reads the zero planes 0 and np as
        # the neighbours of the rows 1 and np-1, as the barrier faces are)
        err = np.max(np.abs(ls[sl] - la[sl])) / np.max(np.abs(la[sl]))
        assert err <= 1e-13, (which, err)


@pytest.mark.parametrize("which", [1, 3])
def test_neumann_closures_annihilate_constants(which):
    """w: constant in r and phi -> only the z second difference (Dirichlet)
    remains; u: constant in z and phi -> only the r rows remain, so L of a
    function of i alone does not depend on j or k."""
    nr, nz, np_ = 7, 6, 9
    s = OS.Sector(nr, nz, np_, **PHYS)
    p = s.p
    i1, j1, k1 = s.ranges(which)
    rng = np.random.default_rng(4)
    psi = np.zeros(s.shape)
    if which == 1:
        f = rng.standard_normal(j1)
        psi[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = f[None, :, None]
        lz = np.zeros(j1)
        for j in range(j1):
            lz[j] = ((f[j + 1] if j + 1 < j1 else 0.0) - 2 * f[j] + (f[j - 1] if j > 0 else 0.0)) / p.dz ** 2
        want = np.broadcast_to(lz[None, :, None], (k1, j1, i1))
    else:
        g = rng.standard_normal(i1)
        psi[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = g[None, None, :]
        ls = _apply(s, which, psi)
        want = np.broadcast_to(ls[1, 1, 1:i1 + 1][None, None, :], (k1, j1, i1))
    got = _apply(s, which, psi)[1:k1 + 1, 1:j1 + 1, 1:i1 + 1]
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), which


@pytest.mark.parametrize("which", [1, 2, 3])
def test_phi_rows_have_closed_form_eigenvectors(which):
    """L (f(i,j) e_m(k)) - e_m(k) L_rz f = f lambda_m e_m / (dp r)^2 with
    e_m = cos(m pi (k - 1/2) / np) (Neumann faces: w, u) or
    e_m = sin(m pi k / np) (Dirichlet faces: v), lambda_m = 2 cos(m pi / n) - 2."""
    nr, nz, np_ = 6, 5, 10
    s = OS.Sector(nr, nz, np_, **PHYS)
    p = s.p
    i1, j1, k1 = s.ranges(which)
    rng = np.random.default_rng(7 + which)
    f = rng.standard_normal((j1, i1))
    k = np.arange(1, k1 + 1)
    r = p.r0 + (np.arange(1, i1 + 1) - (0.0 if which == 3 else 0.5)) * p.dr
    # L_rz f: a plane far from the faces of a field constant in phi
    const = np.zeros(s.shape)
    const[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = f[None]
    lrz = _apply(s, which, const)[k1 // 2, 1:j1 + 1, 1:i1 + 1]
    for m in (1, 2, 5):
        if which == 2:
            e, lam = np.sin(m * np.pi * k / np_), 2 * np.cos(m * np.pi / np_) - 2
        else:
            e, lam = np.cos(m * np.pi * (k - 0.5) / np_), 2 * np.cos(m * np.pi / np_) - 2
        psi = np.zeros(s.shape)
        psi[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = f[None] * e[:, None, None]
        got = _apply(s, which, psi)[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] - e[:, None, None] * lrz[None]
        want = f[None] * lam * e[:, None, None] / (p.dp * r[None, None, :]) ** 2
        assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want)), (which, m)


def test_u_z_rows_have_closed_form_eigenvectors():
    """u's z rows are Neumann: g(i) cos(m pi (j - 1/2)/nz) picks up
    lambda_m / dz^2 over L of g alone."""
    nr, nz, np_ = 6, 8, 6
    s = OS.Sector(nr, nz, np_, **PHYS)
    p = s.p
    i1, j1, k1 = s.ranges(3)
    g = np.random.default_rng(1).standard_normal(i1)
    base = np.zeros(s.shape)
    base[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = g[None, None, :]
    lg = _apply(s, 3, base)[1:k1 + 1, 1:j1 + 1, 1:i1 + 1]
    j = np.arange(1, j1 + 1)
    for m in (1, 3):
        e, lam = np.cos(m * np.pi * (j - 0.5) / nz), 2 * np.cos(m * np.pi / nz) - 2
        psi = np.zeros(s.shape)
        psi[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = g[None, None, :] * e[None, :, None]
        got = _apply(s, 3, psi)[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] - e[None, :, None] * lg
        want = g[None, None, :] * lam * e[None, :, None] / p.dz ** 2
        assert np.max(np.abs(got - want)) <= 1e-11 * np.max(np.abs(want)), m


def test_w_operator_is_self_adjoint_under_r():
    """sum r (a L b) = sum r (b L a): the r rows are the flux form
    (r_{i+1/2} (b_{i+1} - b_i) - r_{i-1/2} (b_i - b_{i-1})) / (r h^2) with
    zero flux through the cylinders, the phi rows symmetric with zero flux
    through the barrier; a wrong closure coefficient or sign breaks this."""
    nr, nz, np_ = 7, 6, 8
    s = OS.Sector(nr, nz, np_, **PHYS)
    p = s.p
    i1, j1, k1 = s.ranges(1)
    rng = np.random.default_rng(11)
    a, b = np.zeros(s.shape), np.zeros(s.shape)
    a[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = rng.standard_normal((k1, j1, i1))
    b[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = rng.standard_normal((k1, j1, i1))
    r = p.r0 + (np.arange(1, i1 + 1) - 0.5) * p.dr
    sl = np.s_[1:k1 + 1, 1:j1 + 1, 1:i1 + 1]
    la, lb = _apply(s, 1, a)[sl], _apply(s, 1, b)[sl]
    x = np.sum(r[None, None, :] * a[sl] * lb)
    y = np.sum(r[None, None, :] * b[sl] * la)
    assert abs(x - y) <= 1e-12 * np.sum(np.abs(r[None, None, :] * a[sl] * lb))


@pytest.mark.parametrize("which", [1, 2, 3])
def test_exact_solve_zeroes_the_residual(which):
    nr, nz, np_ = 6, 5, 8
    s = OS.Sector(nr, nz, np_, **PHYS)
    rng = np.random.default_rng(which)
    rhs = rng.standard_normal(s.shape)
    x = s.poisson_exact(which, rhs)
    m, _ = s.residual(which, x, rhs)
    assert m <= 1e-10 * np.max(np.abs(rhs)), m


def test_poisson_output_conventions():
    """psip <- old psi on the copied range; walls 0; the mirrored ghosts
    -psi of the adjacent unknown (sect360v2mg.f:1439-1457, 1530-1547,
    1630-1654)."""
    nr, nz, np_ = 6, 5, 8
    s = OS.Sector(nr, nz, np_, **PHYS)
    rng = np.random.default_rng(5)
    for n in O3.FIELDS:
        s[n] = rng.standard_normal(s.shape)
    old = {n: s[n].copy() for n in ("wel", "vel", "uel")}
    s.poisson(1)
    w = s["wel"]
    assert np.array_equal(s["welp"][:, :nz + 1, :], old["wel"][:, :nz + 1, :])
    assert np.all(w[:, 0, :] == 0) and np.all(w[:, nz, :] == 0)
    assert np.array_equal(w[1:np_ + 1, 1:nz, 0], -w[1:np_ + 1, 1:nz, 1])
    assert np.array_equal(w[0, 1:nz, 1:nr + 1], -w[1, 1:nz, 1:nr + 1])
    assert np.array_equal(w[np_ + 1, 1:nz, 0], -w[np_, 1:nz, 1])
    s.poisson(2)
    v = s["vel"]
    assert np.array_equal(s["velp"][:np_ + 1], old["vel"][:np_ + 1])
    assert np.all(v[0] == 0) and np.all(v[np_] == 0)
    assert np.all(v[1:np_, :, 0] == 0) and np.all(v[1:np_, :, nr + 1] == 0)
    s.poisson(3)
    u = s["uel"]
    assert np.array_equal(s["uelp"][:, :, :nr + 1], old["uel"][:, :, :nr + 1])
    assert np.all(u[:, :, 0] == 0) and np.all(u[:, :, nr] == 0)
    assert np.array_equal(u[1:np_ + 1, 0, 1:nr], -u[1:np_ + 1, 1, 1:nr])
    assert np.array_equal(u[np_ + 1, 1:nz + 1, 1:nr], -u[np_, 1:nz + 1, 1:nr])
