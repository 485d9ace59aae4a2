"""Pins of the 3D tank oracle (oracle/vvt_oracle.c) that do not depend on the
oracle's own formulas:

* the compact derivative of a line (the rows of velvor2dperiodicrbconv.f:
  613-620 and 1085-1097 solved by its tdma) is exact on quadratics for both
  end-row kinds, agrees with a dense solve of the same rows, and is fourth
  order inside;
* each Poisson operator, on fields held away from the side walls, is
  rot3dmg.f's residual operator (rot3dmg.f:1695-1723, 1927-1965, 2130-2165,
  through the 3D oracle) in the limit of a large radius, where its metric
  terms vanish; along y its end rows have the eigenvectors of mirrored
  (cosine) and zero (sine) walls;
* the explicit step commutes with the mirror images of the tank (x and y
  reflections, the z reflection with gravity reversed), with the vorticity
  transforming as a pseudo-vector and the rotation reversed where the
  mirror reverses it: this fixes the sign and the index offsets of every
  term;
* at rest without buoyancy the vorticity stays zero; the surface heat flux
  of a linear profile is its gradient times the surface area.
"""
import numpy as np
import pytest

import oracle3 as O3
import oracle_tank as OT


def _poly(x, deg, c):
    f = sum(c[p] * x ** p for p in range(deg + 1))
    df = np.zeros_like(x) + sum(p * c[p] * x ** (p - 1) for p in range(1, deg + 1))
    return f, df


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("deg", [0, 1, 2])
def test_compact_line_exact_on_quadratics(kind, deg):
    m, h = 17, 0.13
    t = np.arange(m + 2, dtype=float)
    x = t * h if kind == 0 else np.concatenate([[0.0], (t[1:m + 1] - 0.5) * h, [m * h]])
    c = np.random.default_rng(deg).standard_normal(3)
    f, df = _poly(x, deg, c)
    got = OT.compact_line(kind, h, f)
    # kind 1's end rows carry the reference's truncated decimals
    # (0.533333333, 2.966666667): exact to about 1e-9 of |f|/h
    tol = 1e-11 if kind == 0 else 1e-8 * np.max(np.abs(f)) / h
    assert np.max(np.abs(got - df[1:m + 1])) <= tol * max(1.0, np.max(np.abs(df)))


@pytest.mark.parametrize("kind", [0, 1])
def test_compact_line_is_the_dense_solve_of_its_rows(kind):
    m, h = 12, 0.3
    f = np.random.default_rng(5).standard_normal(m + 2)
    M = np.zeros((m, m))
    for r in range(m):
        M[r, r] = 4.0
        if r > 0:
            M[r, r - 1] = 1.0
        if r < m - 1:
            M[r, r + 1] = 1.0
    rhs = 3.0 * (f[2:] - f[:-2]) / h
    if kind == 0:
        rhs[0] -= (-1.5 * f[0] + 2.0 * f[1] - 0.5 * f[2]) / h
        rhs[-1] -= (0.5 * f[m - 1] - 2.0 * f[m] + 1.5 * f[m + 1]) / h
    else:
        M[0, 0], M[0, 1] = 1.0, 2.0
        M[-1, -2], M[-1, -1] = 2.0, 1.0
        rhs[0] = (0.533333333 * f[0] - 3.6 * f[1] + 2.966666667 * f[2] + 0.1 * f[4]) / h
        rhs[-1] = (-0.1 * f[m - 3] - 2.966666667 * f[m - 1] + 3.6 * f[m] - 0.533333333 * f[m + 1]) / h
    want = np.linalg.solve(M, rhs)
    assert np.allclose(OT.compact_line(kind, h, f), want, rtol=1e-13, atol=1e-13 * np.max(np.abs(want)))


def test_compact_line_fourth_order_inside():
    errs = []
    for m in (32, 64):
        h = 1.0 / (m + 1)
        x = np.arange(m + 2) * h
        got = OT.compact_line(0, h, np.sin(3.0 * x))
        mid = m // 2
        errs.append(abs(got[mid - 1] - 3.0 * np.cos(3.0 * x[mid])))
    assert errs[0] / errs[1] > 12.0, errs


def _tank(nr=6, nz=5, np_=8, Re=37.0, Gr=11.0, **kw):
    return OT.Tank(nr, nz, np_, dt=1e-4, Re=Re, Ptl=5.0, Gr=Gr, **kw)


@pytest.mark.parametrize("which", [1, 2, 3])
def test_operator_is_rot3dmg_operator_at_large_radius(which):
    """Away from the side walls the tank's operator of `which` equals the 3D
    oracle's restatement of rot3dmg.f's residual operator for the same
    velocity component at r0 = 1e7 (metric terms of relative size 1e-7)."""
    nr, nz, np_ = 6, 5, 10
    r0 = 1.0e7
    dr, dz, ly = 0.1, 0.12, 1.3
    t = OT.Tank(nr, nz, np_, dt=1e-4, Re=1.0, Ptl=1.0, Gr=0.0, dr=dr, dz=dz, dp=ly / np_)
    a = O3.Run3(nr, nz, np_, 1e-4, 0.0, 1.0, 1.0, 0.0, r0=r0, rmax=r0 + nr * dr, zmax=nz * dz)
    a.p.dp = (ly / np_) / r0     # arc length dp * r0 = the tank's dy
    rng = np.random.default_rng(which)
    psi_t = np.zeros(t.shape)
    i1, j1, k1 = t.ranges(which)
    psi_t[3:k1 - 1, 1:j1 + 1, 1:i1 + 1] = rng.standard_normal((k1 - 4, j1, i1))
    psi_a = np.zeros(a.shape)
    psi_a[:np_ + 1] = psi_t[:np_ + 1]
    psi_a[0] = psi_a[np_]
    lt = t.apply(which, psi_t)
    _, ra = a.residual(which, psi_a, np.zeros(a.shape))
    la = -ra
    sl = (slice(3, k1 - 1), slice(1, j1 + 1), slice(1, i1 + 1))
    scale = np.max(np.abs(lt[sl]))
    assert np.max(np.abs(lt[sl] - la[sl])) <= 1e-5 * scale


@pytest.mark.parametrize("which,m", [(1, 0), (1, 3), (2, 1), (2, 5), (3, 2)])
def test_operator_side_wall_rows_have_the_cosine_and_sine_modes(which, m):
    t = _tank(np_=9)
    i1, j1, k1 = t.ranges(which)
    npl = t.p.np
    k = np.arange(1, k1 + 1)
    mode = np.sin(np.pi * m * k / npl) if which == 2 else np.cos(np.pi * m * (k - 0.5) / npl)
    lam = (2.0 * np.cos(np.pi * m / npl) - 2.0) / t.p.dp ** 2
    g = np.random.default_rng(m).standard_normal((j1, i1))
    psi = np.zeros(t.shape)
    psi[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = mode[:, None, None] * g[None]
    flat = np.zeros(t.shape)
    flat[1:k1 + 1, 1:j1 + 1, 1:i1 + 1] = g[None]
    lxz = t.apply(which, flat)[1, 1:j1 + 1, 1:i1 + 1] - (t.apply(which, flat)[1, 1:j1 + 1, 1:i1 + 1] -
                                                           t.apply(which, flat)[k1 // 2 + 1, 1:j1 + 1, 1:i1 + 1])
    # L(mode * g) = mode * (L_xz g + lam g)
    want = mode[:, None, None] * (lxz[None] + lam * g[None])
    got = t.apply(which, psi)[1:k1 + 1, 1:j1 + 1, 1:i1 + 1]
    assert np.max(np.abs(got - want)) <= 1e-10 * np.max(np.abs(want))


# ---- mirror images of the explicit step ----

STAG = {  # (x, y, z) staggering: 'c' cells, 'f' faces
    "A": "cff", "B": "fcf", "C": "ffc", "th": "ccc", "uel": "fcc", "vel": "cfc", "wel": "ccf"}
SIGN = {  # sign of each quantity under the reflection of axis x, y, z
    "A": (1, -1, -1), "B": (-1, 1, -1), "C": (-1, -1, 1), "th": (1, 1, 1),
    "uel": (-1, 1, 1), "vel": (1, -1, 1), "wel": (1, 1, -1)}
BASE = {n: n for n in STAG}
BASE.update({"An": "A", "Ap": "A", "Bn": "B", "Bp": "B", "Cn": "C", "Cp": "C", "thn": "th", "thp": "th",
             "uelp": "uel", "velp": "vel", "welp": "wel"})


def _reflect(a, name, axis, n):
    """mirror image of field `name` along axis 0 (x), 1 (y), 2 (z); n is
    (nr, np, nz)"""
    b = BASE[name]
    st, sg = STAG[b][axis], SIGN[b][axis]
    ax = {0: 2, 1: 0, 2: 1}[axis]
    size = a.shape[ax]
    out = np.zeros_like(a)
    nn = n[axis]
    last = nn + 1 if st == "c" else nn
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    src[ax] = slice(0, last + 1)
    dst[ax] = slice(last, None, -1) if last + 1 == size else slice(last, None if last + 1 == size else None, -1)
    idx = np.arange(last + 1)
    dst_idx = last - idx
    moved = np.take(a, idx, axis=ax)
    out_view = np.moveaxis(out, ax, 0)
    out_view[dst_idx] = np.moveaxis(moved, ax, 0)
    return sg * out


def _random_state(t, seed):
    rng = np.random.default_rng(seed)
    for nme in ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "uel", "vel", "wel"):
        t[nme] = rng.standard_normal(t.shape)
    # quantities have no values beyond their last face: keep those planes 0
    n = (t.p.nr, t.p.np, t.p.nz)
    for nme in ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "uel", "vel", "wel"):
        t[nme] = _reflect(_reflect(t[nme], nme, 0, n), nme, 0, n) * 1.0
        t[nme] = _reflect(_reflect(t[nme], nme, 1, n), nme, 1, n) * 1.0
        t[nme] = _reflect(_reflect(t[nme], nme, 2, n), nme, 2, n) * 1.0


def _explicit(t):
    t.derivatives()
    for rt in ("avorticity", "bvorticity", "cvorticity", "temperature"):
        t.call(rt)
    return {n: t[n].copy() for n in ("An", "Bn", "Cn", "thn")}


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_explicit_step_commutes_with_mirror_images(axis):
    nr, nz, np_ = 7, 6, 9
    Re, Gr = 37.0, 11.0
    t1 = _tank(nr, nz, np_, Re=Re, Gr=Gr)
    _random_state(t1, 3)
    # reversed rotation under the x and y mirrors, reversed gravity under z
    t2 = _tank(nr, nz, np_, Re=-Re if axis < 2 else Re, Gr=-Gr if axis == 2 else Gr)
    n = (nr, np_, nz)
    for nme in ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "uel", "vel", "wel"):
        t2[nme] = _reflect(t1[nme], nme, axis, n)
    out1 = _explicit(t1)
    out2 = _explicit(t2)
    for nme in ("An", "Bn", "Cn", "thn"):
        want = _reflect(out1[nme], nme, axis, n)
        scale = np.max(np.abs(want))
        assert scale > 0
        assert np.max(np.abs(out2[nme] - want)) <= 1e-11 * scale, nme


def test_update_commutes_with_the_x_mirror():
    nr, nz, np_ = 7, 6, 9
    t1 = _tank(nr, nz, np_)
    _random_state(t1, 4)
    t2 = _tank(nr, nz, np_, Re=-37.0)
    n = (nr, np_, nz)
    for nme in ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "uel", "vel", "wel"):
        t2[nme] = _reflect(t1[nme], nme, 0, n)
    for t in (t1, t2):
        _explicit(t)
        t.call("update")
    for nme in ("A", "B", "C", "th", "Ap", "thp"):
        want = _reflect(t1[nme], nme, 0, n)
        assert np.max(np.abs(t2[nme] - want)) <= 1e-11 * np.max(np.abs(want)), nme


def test_rest_without_buoyancy_stays_at_rest():
    t = _tank(Gr=0.0)
    rng = np.random.default_rng(2)
    th = rng.standard_normal(t.shape)
    t["th"], t["thp"] = th, th
    for _ in range(2):
        t.step()
    for nme in ("A", "B", "C", "uel", "vel", "wel"):
        assert np.all(t[nme] == 0.0), nme


def test_heat_flux_of_a_linear_profile():
    nr, nz, np_ = 6, 9, 8
    t = _tank(nr, nz, np_)
    z = np.concatenate([[0.0], (np.arange(1, nz + 1) - 0.5) * t.p.dz, [nz * t.p.dz]])
    s = 0.7
    t["th"] = np.broadcast_to((1.0 + s * z)[None, :, None], t.shape)
    trans1, trans2 = t.heat()
    area = nr * t.p.dr * (np_ // 2) * t.p.dp
    # the routine's weights are truncated decimals of Fornberg's (-352/105, ...)
    assert abs(trans1 - s * area) <= 1e-6 * s * area
    assert abs(trans2 + s * area) <= 1e-6 * s * area


def test_oracle_solve_zeroes_the_residual():
    t = _tank(6, 5, 8)
    rng = np.random.default_rng(1)
    for which in (1, 2, 3):
        f = rng.standard_normal(t.shape)
        psi = t.solve(which, f)
        m, _ = t.residual(which, psi, f)
        assert m <= 1e-11 * np.max(np.abs(f)) * 100


@pytest.mark.parametrize("shape", [(8, 6, 10), (7, 9, 6)])
def test_fast_solve_is_the_dense_solve(shape):
    """The separable solve the benchmark's CPU leg uses at full size (DCT-II
    / DST-I along y, eigenbases along z and x) equals the dense solve."""
    t = _tank(*shape)
    fs = OT.FastSolver(t)
    rng = np.random.default_rng(7)
    for which in (1, 2, 3):
        f = rng.standard_normal(t.shape)
        a, b = t.solve(which, f), fs.solve(which, f)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
