"""Writes the golden fixtures of tests/golden/ from the reference build.

Run in the container that holds the reference sources, after
`make -C oracle all` has compiled velvor2dperiodicrbconv.f into
oracle/_ref/librb.so:

    python tests/golden/make_golden.py

Every fixture holds inputs and the reference's outputs only (numpy .npz):
  banded.npz         tdma / ptdma / pdma (rb.f:3046-3164) on seeded systems
  poisson_16x32.npz  poisson2 / poisson3 (rb.f:1796-2059) as shipped
                     (absolute stop 2e-6) on a seeded right-hand side
  traj_16x32.npz     the time loop (rb.f:341-382) from the seeded
                     temperature, 20 steps: fields after steps 1, 2, 20 and
                     the wall fluxes trans1/trans2 of every step
  traj_64x128.npz    the same at 64x128 for 1000 steps (fields after steps
                     1 and 1000, fluxes of every step)
  ann_16x32.npz      the rotating-annulus plane (velvor2dnewsch3.4.f built
                     into oracle/_ref/libann.so), r0=0.5, rmax=1, Ra=1e5,
                     Re=100, Pr=1, dt=1e-5, 20 steps (fields after 1, 20)
  ann_64x128.npz     the same at 64x128, Ra=1e7, Re=sqrt(1e5), dt=1e-6,
                     300 steps (fields after 1, 300)
The time-loop runs use box width 3 (the reference's hard-coded xx,
rb.f:132), r0=0.5, rmax=1, Ra=1e5, Pr=1, dt=1e-5, and the temperature
seed of the oracle (splitmix64, seed 20261202, amplitude 1e-3).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import refdrv  # noqa: E402
import oracle as O  # noqa: E402


def banded(ref):
    import ctypes
    rng = np.random.default_rng(11)
    out = {}
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for n in (7, 33, 129):
        for kind in ("tdma", "ptdma", "pdma"):
            a = rng.uniform(-1, 1, n)
            b = rng.uniform(3, 5, n)
            c = rng.uniform(-1, 1, n)
            d = rng.uniform(-0.5, 0.5, n)
            e = rng.uniform(-0.3, 0.3, n)
            r = rng.uniform(-1, 1, n)
            if kind == "pdma":
                b = rng.uniform(-1, 1, n)
                c = rng.uniform(4, 6, n)
                a[:2] = 0.0
                b[0] = 0.0
                d[-1] = 0.0
                e[-2:] = 0.0
            ins = {"a": a, "b": b, "c": c, "d": d, "e": e, "r": r}
            w = {k: v.copy() for k, v in ins.items()}
            y = np.zeros(n)
            nn = ctypes.c_int(n)
            if kind == "tdma":
                ref.L.tdma_(P(w["a"]), P(w["b"]), P(w["c"]), P(y), P(w["r"]), ctypes.byref(nn))
            elif kind == "ptdma":
                ref.L.ptdma_(P(w["a"]), P(w["b"]), P(w["c"]), P(y), P(w["r"]), ctypes.byref(nn))
            else:
                ref.L.pdma_(P(w["a"]), P(w["b"]), P(w["c"]), P(w["d"]), P(w["e"]), P(w["r"]),
                            P(y), ctypes.byref(nn))
            for k, v in ins.items():
                out[f"{kind}_{n}_{k}"] = v
            out[f"{kind}_{n}_y"] = y
    np.savez_compressed(os.path.join(HERE, "banded.npz"), **out)


def poisson(ref, nr=16, np_=32):
    ref.setup(nr, np_, 1e-5, 1e5, 1.0)
    rng = np.random.default_rng(7)
    out = {"nr": nr, "np": np_, "dr": ref.dr, "dp": ref.dp}
    for which in (2, 3):
        ni = nr + 2 if which == 2 else nr + 1
        k = np.arange(np_ + 1)[:, None]
        i = np.arange(ni)[None, :]
        rhs = np.sin(2 * np.pi * k / np_) * np.cos(0.3 * i) + 0.1 * rng.standard_normal((np_ + 1, ni))
        rhs[np_] = rhs[0]
        psi0 = 1e-3 * rng.standard_normal((np_ + 1, ni))
        psi0[:, 0] = 0.0
        psi0[:, ni - 1] = 0.0
        psi0[np_] = psi0[0]
        psip0 = 0.5 * psi0
        big = lambda a: np.pad(a, ((0, refdrv.KP1 - a.shape[0]), (0, refdrv.KR1 - a.shape[1])))
        psi, psip, r = big(psi0), big(psip0), big(rhs)
        ref.poisson(which, psi, psip, r)
        out[f"p{which}_rhs"] = rhs
        out[f"p{which}_psi_in"] = psi0
        out[f"p{which}_psip_in"] = psip0
        out[f"p{which}_psi"] = psi[: np_ + 1, :ni].copy()
        out[f"p{which}_psip"] = psip[: np_ + 1, :ni].copy()
    np.savez_compressed(os.path.join(HERE, f"poisson_{nr}x{np_}.npz"), **out)


def trajectory(ref, nr, np_, nsteps, snaps):
    dt, Ra, Pr = 1e-5, 1e5, 1.0
    run = O.Run(nr, np_, dt, Ra, Pr, width=3.0)
    run.seed()
    th0 = run["th"].copy()
    ref.setup(nr, np_, dt, Ra, Pr, width=3.0)
    ref.load("th", th0)
    ref.start()
    out = {"nr": nr, "np": np_, "dt": dt, "Ra": Ra, "Pr": Pr, "width": 3.0,
           "r0": 0.5, "rmax": 1.0, "seed": 20261202, "amp": 1e-3, "th0": th0}
    trans = np.zeros((nsteps, 2))
    for s in range(1, nsteps + 1):
        trans[s - 1] = ref.step()
        if s in snaps:
            for name, ni in (("th", nr + 2), ("thp", nr + 2), ("C", nr + 1), ("Cp", nr + 1),
                             ("uel", nr + 1), ("vel", nr + 2)):
                out[f"s{s}_{name}"] = ref.get(name, ni)
    out["trans"] = trans
    np.savez_compressed(os.path.join(HERE, f"traj_{nr}x{np_}.npz"), **out)


def annulus(nr, np_, nsteps, snaps, dt, Ra, Re):
    ref = refdrv.Reference("ann")
    Pr = 1.0
    run = O.Run(nr, np_, dt, Ra, Pr, geom=1, r0=0.5, Re=Re)
    run.seed()
    th0 = run["th"].copy()
    ref.setup(nr, np_, dt, Ra, Pr, Re=Re)
    ref.load("th", th0)
    ref.start()
    out = {"nr": nr, "np": np_, "dt": dt, "Ra": Ra, "Pr": Pr, "Re": Re, "r0": 0.5, "rmax": 1.0,
           "seed": 20261202, "amp": 1e-3, "th0": th0}
    trans = np.zeros((nsteps, 2))
    for s in range(1, nsteps + 1):
        trans[s - 1] = ref.step()
        if s in snaps:
            for name, ni in (("th", nr + 2), ("thp", nr + 2), ("C", nr + 1), ("Cp", nr + 1),
                             ("uel", nr + 1), ("vel", nr + 2)):
                out[f"s{s}_{name}"] = ref.get(name, ni)
    out["trans"] = trans
    np.savez_compressed(os.path.join(HERE, f"ann_{nr}x{np_}.npz"), **out)


def main():
    if not refdrv.available():
        raise SystemExit("oracle/_ref/librb.so missing: run make -C oracle all first")
    ref = refdrv.Reference()
    banded(ref)
    poisson(ref)
    trajectory(ref, 16, 32, 20, {1, 2, 20})
    trajectory(ref, 64, 128, 1000, {1, 1000})
    if refdrv.available("ann"):
        annulus(16, 32, 20, {1, 20}, 1e-5, 1e5, 100.0)
        annulus(64, 128, 300, {1, 300}, 1e-6, 1e7, float(np.sqrt(1e5)))
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
