"""Writes tests/golden/r3_*.npz from the 3D reference build
(oracle/_ref/librot3d.so = rot3dmg.f compiled unmodified; only in the
container that holds the reference sources).  The fixtures hold inputs and
the reference's outputs only.

  r3_explicit.npz: a seeded random state (velocities scaled so that every
    hybrid-upwind branch of rot3dmg.f fires) and the reference's outputs of
    Avorticity, Bvorticity, Cvorticity, temperature, heat, rhsw, rhsv, rhsu
    and update applied in that order (rot3dmg.f:240-253 without the solves).
  r3_resid.npz: for each Poisson operator, a seeded iterate and right-hand
    side, the iterate after one relaxation pass and the residual the
    reference stores for it (relax1/2/3 on the finest level).
  r3h_explicit.npz: as r3_explicit.npz for rot3dmghor.f
    (oracle/_ref/librot3dhor.so, the stationary horizontal annulus), from
    its own seeded state; "heat" holds (trans1, trans2).

    python tests/golden/make_golden3.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle")]
import refdrv3 as R3  # noqa: E402
import oracle3 as O3  # noqa: E402

NR, NZ, NP = 16, 12, 24
PARAMS = dict(dt=1e-4, Ra=1e4, Re=30.0, Ptl=1.0, Gr=5.0)
VSCALE = 300.0
SEED = 20261204


def state(rng):
    K = NP + 1
    st = {}
    for n in O3.FIELDS:
        scale = VSCALE if n in ("uel", "vel", "wel", "uelp", "velp", "welp") else 1.0
        a = scale * rng.standard_normal((K, NZ + 2, NR + 2))
        a[0] = a[NP]
        st[n] = a
    return st


def main():
    ref = R3.Reference3()
    run = O3.Run3(NR, NZ, NP, **PARAMS)
    ref.setup(run.p)
    rng = np.random.default_rng(SEED)
    st = state(rng)
    K = NP + 1
    sl = (slice(0, K), slice(0, NZ + 2), slice(0, NR + 2))
    for n, a in st.items():
        ref.f[n][...] = 0.0
        ref.f[n][sl] = a
    out = {"in_" + n: a for n, a in st.items()}
    for rt, name in [("avorticity", "An"), ("bvorticity", "Bn"), ("cvorticity", "Cn"), ("temperature", "thn")]:
        ref.call(rt)
        out[rt] = ref.f[name][sl].copy()
    out["heat"] = np.array(ref.heat())
    for w, name in [(1, "wrh"), (2, "vrh"), (3, "urh")]:
        ref.rhs(w)
        out[name] = ref.f[name][sl].copy()
    ref.call("update")
    for n in ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "An", "Bn", "Cn", "thn"):
        out["upd_" + n] = ref.f[n][sl].copy()
    np.savez_compressed(os.path.join(HERE, "r3_explicit.npz"), **out)
    res = {}
    for w in (1, 2, 3):
        psi = rng.standard_normal((K, NZ + 2, NR + 2))
        rhs = rng.standard_normal((K, NZ + 2, NR + 2))
        psi[0] = psi[NP]
        rhs[0] = rhs[NP]
        fp = np.zeros(R3.SHAPE)
        fr = np.zeros(R3.SHAPE)
        fp[sl] = psi
        fr[sl] = rhs
        psif, rc = ref.relax_residual(w, run.p, fp, fr)
        res[f"psi{w}"] = psi
        res[f"rhs{w}"] = rhs
        res[f"psif{w}"] = psif[sl]
        res[f"rc{w}"] = rc[sl]
    np.savez_compressed(os.path.join(HERE, "r3_resid.npz"), **res)
    hor_fixture()
    print("wrote r3_explicit.npz, r3_resid.npz, r3h_explicit.npz")


def hor_fixture():
    ref = R3.Reference3(1)
    run = O3.Run3(NR, NZ, NP, model=1, **PARAMS)
    ref.setup(run.p)
    st = state(np.random.default_rng(SEED + 1))
    K = NP + 1
    sl = (slice(0, K), slice(0, NZ + 2), slice(0, NR + 2))
    for n, a in st.items():
        ref.f[n][...] = 0.0
        ref.f[n][sl] = a
    out = {"in_" + n: a for n, a in st.items()}
    for rt, name in [("avorticity", "An"), ("bvorticity", "Bn"), ("cvorticity", "Cn"), ("temperature", "thn")]:
        ref.call(rt)
        out[rt] = ref.f[name][sl].copy()
    out["heat"] = np.array(ref.heat2())
    for w, name in [(1, "wrh"), (2, "vrh"), (3, "urh")]:
        ref.rhs(w)
        out[name] = ref.f[name][sl].copy()
    ref.call("update")
    for n in ("A", "Ap", "B", "Bp", "C", "Cp", "th", "thp", "An", "Bn", "Cn", "thn"):
        out["upd_" + n] = ref.f[n][sl].copy()
    np.savez_compressed(os.path.join(HERE, "r3h_explicit.npz"), **out)


if __name__ == "__main__":
    main()
