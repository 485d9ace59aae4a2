"""Pins of the meridional-plane oracle (oracle/vvm_oracle.c) that do not depend
on the oracle's own formulas:

* Fornberg's weights (wfdif.f:59-85) reproduce the textbook three-point
  weights on uniform and uneven nodes and equal the solution of the
  Vandermonde (Taylor) system on random nodes;
* the tanh nodes are uniform at beta = 0, mirror-symmetric and clustered
  toward the walls;
* the compact derivative is exact on polynomials up to degree 4 on stretched
  nodes (not 5), converges at fourth order, and on uniform nodes is the
  classical Pade scheme (1/4, 1, 1/4; 3/(4h)) with the classical one-sided
  fourth-order end derivatives;
* on fields that are polynomials of degree two in r and in z, where every
  discrete derivative of the step is exact, the DuFort-Frankel update equals
  the leapfrog of the continuous equations written in vvm_oracle.h (every
  term, sign and factor, from analytic derivatives), the right-hand sides are
  omega_z and -omega_r - omega/r, the wall vorticity is -w_r and u_z, the
  operators are lap - 1/r^2 and lap, and the exact solve returns the
  manufactured field;
* the separable solve equals a sparse direct solve; the step commutes with
  the mirror z -> H - z (vorticity and axial velocity odd); without forcing
  the kinetic energy decays.
"""
import numpy as np
import pytest

import oracle_merid as OM

R0, RMAX, H = 0.5, 1.0, 0.5


# ---------------------------------------------------------------- Fornberg
def test_fornberg_uniform_three_point():
    c = OM.fornberg(np.array([2.0, 2.5, 3.0]), 2.5, 2)
    assert np.allclose(c[0], [0, 1, 0], atol=0)
    assert np.allclose(c[1], [-1.0, 0.0, 1.0], rtol=1e-15, atol=1e-15)
    assert np.allclose(c[2], [4.0, -8.0, 4.0], rtol=1e-15)


def test_fornberg_uneven_three_point():
    hm, hp = 0.3, 0.7
    c = OM.fornberg(np.array([-hm, 0.0, hp]), 0.0, 2)
    s = hm + hp
    assert np.allclose(c[1], [-hp / (hm * s), (hp - hm) / (hm * hp), hm / (hp * s)], rtol=1e-14)
    assert np.allclose(c[2], [2 / (hm * s), -2 / (hm * hp), 2 / (hp * s)], rtol=1e-14)


@pytest.mark.parametrize("n,m", [(3, 2), (5, 1), (5, 4), (7, 3)])
def test_fornberg_is_the_taylor_system(n, m):
    rng = np.random.default_rng(n * 10 + m)
    x = np.sort(rng.uniform(0.0, 1.0, n))
    xx = rng.uniform(0.0, 1.0)
    c = OM.fornberg(x, xx, m)
    V = np.array([(x - xx) ** p for p in range(n)])          # V[p, q]
    import math
    for k in range(m + 1):
        e = np.zeros(n)
        e[k] = math.factorial(k)
        ref = np.linalg.solve(V, e)
        assert np.allclose(c[k], ref, rtol=1e-8, atol=1e-8 * np.abs(ref).max())


# ---------------------------------------------------------------- nodes
def test_nodes():
    assert np.allclose(OM.nodes(16, 0.5, 1.0, 0.0), np.linspace(0.5, 1.0, 17), rtol=0, atol=1e-15)
    x = OM.nodes(64, 0.5, 1.0, 1.5)
    assert x[0] == 0.5 and x[-1] == 1.0
    assert np.all(np.diff(x) > 0)
    assert np.allclose(x + x[::-1], 1.5, atol=1e-15)
    h = np.diff(x)
    assert h[0] < 0.4 * h[32] and h[-1] < 0.4 * h[31]


# ---------------------------------------------------------------- compact
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_compact_exact_to_degree_four(deg):
    x = OM.nodes(40, R0, RMAX, 1.5)
    c = np.random.default_rng(deg).standard_normal(deg + 1)
    f = sum(c[p] * x ** p for p in range(deg + 1)) + 0 * x
    df = sum(p * c[p] * x ** (p - 1) for p in range(1, deg + 1)) + 0 * x
    got = OM.compact_line(x, f)
    assert np.max(np.abs(got - df[1:-1])) <= 1e-11 * max(1.0, np.abs(df).max())


def test_compact_not_exact_at_degree_five():
    x = OM.nodes(40, R0, RMAX, 1.5)
    got = OM.compact_line(x, x ** 5)
    assert np.max(np.abs(got - 5 * x[1:-1] ** 4)) > 1e-9


def test_compact_fourth_order():
    errs = []
    for n in (32, 64, 128):
        x = OM.nodes(n, R0, RMAX, 1.5)
        errs.append(np.max(np.abs(OM.compact_line(x, np.sin(7 * x)) - 7 * np.cos(7 * x[1:-1]))))
    assert errs[0] / errs[1] > 12 and errs[1] / errs[2] > 12


def test_compact_uniform_is_classical_pade():
    n, h = 20, 0.05
    x = np.arange(n + 1) * h
    f = np.random.default_rng(3).standard_normal(n + 1)
    d0 = (-25 * f[0] + 48 * f[1] - 36 * f[2] + 16 * f[3] - 3 * f[4]) / (12 * h)
    dn = (25 * f[n] - 48 * f[n - 1] + 36 * f[n - 2] - 16 * f[n - 3] + 3 * f[n - 4]) / (12 * h)
    M = np.zeros((n - 1, n - 1))
    rhs = 0.75 * (f[2:] - f[:-2]) / h
    for i in range(n - 1):
        M[i, i] = 1.0
        if i > 0:
            M[i, i - 1] = 0.25
        if i < n - 2:
            M[i, i + 1] = 0.25
    rhs[0] -= 0.25 * d0
    rhs[-1] -= 0.25 * dn
    ref = np.linalg.solve(M, rhs)
    assert np.allclose(OM.compact_line(x, f), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


# ---------------------------------------------------------------- polynomial fields
class Poly:
    """p(r, z) = sum c[a, b] r^a z^b, a, b <= 2, with analytic derivatives."""

    def __init__(self, c):
        self.c = np.asarray(c, dtype=float)

    def __call__(self, r, z, dr=0, dz=0):
        out = 0.0
        for a in range(3):
            for b in range(3):
                if a < dr or b < dz:
                    continue
                fa = np.prod(range(a - dr + 1, a + 1)) if dr else 1.0
                fb = np.prod(range(b - dz + 1, b + 1)) if dz else 1.0
                out = out + self.c[a, b] * fa * fb * r ** (a - dr) * z ** (b - dz)
        return out

    def lap(self, r, z):
        return self(r, z, 2, 0) + self(r, z, 1, 0) / r + self(r, z, 0, 2)


def _grid(m):
    return np.meshgrid(m.r, m.z)   # (nz+1, nr+1): R[j, i] = r_i


def _centre(m):
    """Centre coefficient of the three-point r and z second derivatives."""
    d2r = np.array([OM.three_point(m.r, i)[1][1] for i in range(1, m.nr)])
    d2z = np.array([OM.three_point(m.z, j)[1][1] for j in range(1, m.nz)])
    return d2z[:, None] + d2r[None, :]


def test_explicit_is_the_leapfrog_of_the_equations():
    nr, nz, dt, Ra, Re, Pr = 24, 16, 1e-4, 3e3, 7.0, 0.7
    m = OM.Merid(nr, nz, dt, Ra, Re, Pr)
    R, Z = _grid(m)
    rng = np.random.default_rng(11)
    P = {k: Poly(rng.standard_normal((3, 3))) for k in ("w", "v", "th", "u", "ww")}
    for k in P:
        m.f[k][...] = P[k](R, Z)
    m.f["wp"][...] = m.f["w"]
    m.f["vp"][...] = m.f["v"]
    m.f["thp"][...] = m.f["th"]
    m.derivatives()
    m.explicit()
    s = (slice(1, -1), slice(1, -1))
    r, z = R[s], Z[s]
    w, v, th, u, ww = (P[k](r, z) for k in ("w", "v", "th", "u", "ww"))
    RRP = Ra / Re / Pr
    wr, wz = P["w"](r, z, 1, 0), P["w"](r, z, 0, 1)
    vr, vz = P["v"](r, z, 1, 0), P["v"](r, z, 0, 1)
    tr, tz = P["th"](r, z, 1, 0), P["th"](r, z, 0, 1)
    # (1/Pr) w_t = Rw, (1/Pr) v_t = Rv, th_t = Rt  (the equations of vvm_oracle.h)
    Rw = (P["w"].lap(r, z) - w / r ** 2 + 2 / Pr * v * vz / r + 2 * Re * vz - Ra * r * tz
          - 2 * RRP * (v * tz + th * vz) - (u * wr + ww * wz - u * w / r) / Pr)
    Rv = P["v"].lap(r, z) - v / r ** 2 - 2 * Re * u + 2 * RRP * th * u - (u * vr + ww * vz + u * v / r) / Pr
    Rt = P["th"].lap(r, z) - (u * tr + ww * tz)
    # leapfrog with the centre term of the diffusion averaged over the two
    # levels (previous = current here): f_new = f + R / (1/(2 P dt) - D/2)
    Dc = _centre(m)
    for name, Rf, D, Pf in (("wn", Rw, Dc - 1 / r ** 2, Pr), ("vn", Rv, Dc - 1 / r ** 2, Pr),
                            ("thn", Rt, Dc, 1.0)):
        f0 = m.f[name[:-1]][s]
        want = f0 + Rf / (1 / (2 * Pf * dt) - D / 2)
        scale = np.abs(Rf).max() / np.abs(1 / (2 * Pf * dt) - D / 2).min() + np.abs(f0).max() * 1e-3
        assert np.max(np.abs(m.f[name][s] - want)) <= 1e-9 * scale, name


def test_rhs_and_operators_on_quadratics():
    m = OM.Merid(20, 12, 1e-5, 1e3, 5.0, 1.0)
    R, Z = _grid(m)
    p = Poly(np.random.default_rng(2).standard_normal((3, 3)))
    m.f["wn"][...] = p(R, Z)
    m.rhs()
    s = (slice(1, -1), slice(1, -1))
    r, z = R[s], Z[s]
    sc = np.abs(p(r, z, 1, 0)).max() + np.abs(p(r, z, 0, 1)).max() + np.abs(p(r, z)).max() / R0
    assert np.max(np.abs(m.f["fu"][s] - p(r, z, 0, 1))) <= 1e-10 * sc
    assert np.max(np.abs(m.f["fw"][s] - (-p(r, z, 1, 0) - p(r, z) / r))) <= 1e-10 * sc
    # manufactured field zero on the walls: L exact, the solve returns it
    f = (R - R0) * (RMAX - R) * Z * (H - Z)
    q = Poly(np.outer([-R0 * RMAX, R0 + RMAX, -1.0], [0.0, H, -1.0]))
    assert np.allclose(q(R, Z), f, atol=1e-15)
    for which in (1, 2):
        Lf = q.lap(R, Z) - (f / R ** 2 if which == 1 else 0.0)
        got = m.apply(which, f)
        assert np.max(np.abs(got[s] - Lf[s])) <= 1e-10 * np.abs(Lf[s]).max()
        rhs = np.zeros(m.shape)
        rhs[s] = Lf[s]
        sol = m.solve(which, rhs)
        assert np.max(np.abs(sol - f)) <= 1e-11 * np.abs(f).max()


def test_update_walls():
    m = OM.Merid(20, 12, 1e-5, 1e3, 5.0, 1.0)
    R, Z = _grid(m)
    m.f["ww"][...] = (R - R0) * (RMAX - R) * (1 + Z)          # axial velocity zero on the cylinders
    m.f["u"][...] = Z * (H - Z) * (2 + R)                     # radial velocity zero on the lids
    m.f["thn"][...] = 3.0 + Z ** 2                            # insulated bottom
    m.f["vn"][...] = 1.0
    w_old = np.random.default_rng(4).standard_normal(m.shape)
    m.f["w"][...] = w_old
    m.update()
    z = m.z[1:-1]
    assert np.allclose(m.f["w"][1:-1, 0], -(RMAX - R0) * (1 + z), rtol=1e-12)
    assert np.allclose(m.f["w"][1:-1, -1], (RMAX - R0) * (1 + z), rtol=1e-12)
    r = m.r[1:-1]
    assert np.allclose(m.f["w"][0, 1:-1], H * (2 + r), rtol=1e-12)
    assert np.allclose(m.f["w"][-1, 1:-1], -H * (2 + r), rtol=1e-12)
    assert m.f["w"][0, 0] == 0 and m.f["w"][-1, -1] == 0
    assert np.allclose(m.f["th"][0, 1:-1], 3.0, rtol=1e-13)
    assert np.all(m.f["th"][:, 0] == 0.0) and np.all(m.f["th"][:, -1] == 1.0)
    assert np.all(m.f["v"][0] == 0) and np.all(m.f["v"][:, 0] == 0) and np.all(m.f["v"][:, -1] == 0)
    assert np.array_equal(m.f["wp"], w_old)
    # insulated top
    m.f["thn"][...] = 3.0 + (Z - H) ** 2
    m.update()
    assert np.allclose(m.f["th"][-1, 1:-1], 3.0, rtol=1e-13)


def test_heat_of_a_linear_profile():
    m = OM.Merid(30, 14, 1e-5, 1e3, 5.0, 1.0)
    R, Z = _grid(m)
    m.f["th"][...] = 0.2 + 1.7 * R
    assert np.allclose(m.heat(), [1.7 * H, 1.7 * H], rtol=1e-12)


# ---------------------------------------------------------------- solves and trajectories
@pytest.mark.parametrize("which", [1, 2])
def test_separable_solve_is_the_direct_solve(which):
    m = OM.Merid(40, 24, 1e-5, 1e3, 5.0, 1.0)
    rhs = np.zeros(m.shape)
    rhs[1:-1, 1:-1] = np.random.default_rng(which).standard_normal((23, 39))
    a, b = m.solve(which, rhs), m.solve_sparse(which, rhs)
    assert np.max(np.abs(a - b)) <= 1e-12 * np.abs(b).max()
    assert m.residual(which, a, rhs) <= 1e-13


def test_step_commutes_with_the_z_mirror():
    m = OM.Merid(32, 16, 5e-7, 1e6, 30.0, 1.0)
    OM.initial_state(m, 7, amp=1e-2)
    th = m.f["th"]
    th[...] = 0.5 * (th + th[::-1])                     # even in z about H/2
    m.f["thp"][...] = th
    for _ in range(30):
        m.step()
    for k, sign in (("w", -1), ("ww", -1), ("u", 1), ("v", 1), ("th", 1)):
        a = m.f[k]
        sc = max(np.abs(a).max(), 1e-300)
        assert np.max(np.abs(a - sign * a[::-1])) <= 1e-9 * sc, k
    assert np.abs(m.f["w"]).max() > 0 and np.abs(m.f["v"]).max() > 0


def test_energy_decays_without_forcing():
    m = OM.Merid(24, 12, 1e-5, 0.0, 0.0, 1.0)
    R, Z = _grid(m)
    m.f["w"][...] = np.sin(2 * np.pi * (R - R0)) * np.sin(4 * np.pi * Z)
    m.f["wp"][...] = m.f["w"]
    m.f["v"][...] = np.sin(2 * np.pi * (R - R0)) * np.sin(2 * np.pi * Z)
    m.f["vp"][...] = m.f["v"]
    e = []
    for _ in range(40):
        m.step()
        e.append(m.energy())
    # the leapfrog's two time levels decay separately
    assert all(b < a for a, b in zip(e[0::2], e[2::2]))
    assert all(b < a for a, b in zip(e[1::2], e[3::2]))
    assert e[-1] < 0.99 * e[0]
