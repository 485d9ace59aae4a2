"""Benchmark of the velocity-vorticity time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--no-extra]
    python bench.py --rot3d-only [--no-cpu]      # the 3D annulus step (rot3dmg.f) alone
    python bench.py --serve-only                 # the served mixed-size sweep alone
    python bench.py --merid-only [--no-cpu]      # the meridional plane's 5000-step run alone
    python bench.py --sector-only                # the barrier sector (sect360v2mg.f) alone

The default workload is the largest single-GPU one: the 3D
horizontal-convection tank of 512 x 256 x 256 intervals (aspect 4:1:1),
Ra = 1e9, Pr = 5, Ta = 1e8, uniform temperature plus seeded noise.  A step
is one pass of its time loop (compact derivatives along the three axes,
A/B/C and temperature updates, heat fluxes, the three exact Poisson solves,
the wall update) over the device-resident fields.  The timed region runs K
steps after W warm-up steps; inputs are resident in HBM before it starts.  A
second pass of K steps with HIP events around every launch gives the
dominant kernel's average duration and its achieved algorithmic bandwidth
(or, for a transform GEMM, flop rate).  The CPU baseline (rank 0, N=1 only)
times the oracle's step of the same workload on this host.  Unless
--no-extra is given, the line also carries the meridional plane with swirl
on stretched nodes (its whole 5000-step run at 513 x 257), the 2D
legs (the rotating-annulus plane at 513x257 with its V-cycle alternative, a
large V-cycle), the 3D annulus of rot3dmg.f and the served sweep.

The 2D workloads (--workload box-129x65, annulus-513x257, ...) keep their
own headline form; every 2D leg refuses a run whose Poisson solves did not
reach the tolerance (status != 0) and reports the final relative residuals.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "time-steps/s & V-cycle HBM GB/s; served solves/s + p99 at 1e-10 residual"

# name -> (description, config kwargs (geom 1 = annulus), seed, noise amplitude)
WORKLOADS = {
    "annulus-513x257": (
        "rotating-annulus velocity-vorticity plane (velvor2dnewsch3.4.f physics: cylindrical "
        "terms, centrifugal buoyancy, compact 4th-order scheme, DuFort-Frankel), 513x257 nodes "
        "(nr=512 radial, np=256 azimuthal), radius ratio 0.5, Ra=1e8, Ta=1e5 (Re=sqrt(Ta)), "
        "Pr=1, dt=4e-8, T linear in r + U(-1e-4,1e-4); the 513x257 grid of the meridional-plane "
        "annulus, whose swirl/tanh-stretched form has no reference program and is not built",
        dict(geom=1, nr=512, np=256, dt=4e-8, Ra=1e8, Re=1e5 ** 0.5, Pr=1.0, robert=0.05),
        20261203, 1e-4),
    "box-129x65": ("periodic Rayleigh-Benard box (velvor2dperiodicrbconv.f), 129x65 nodes, 2:1, "
                   "Ra=1e5, Pr=1, dt=1e-5 (the reference's CPU-runnable case)",
                   dict(geom=0, nr=64, np=128, width=1.0, dt=1e-5, Ra=1e5, Pr=1.0), 20261202, 1e-3),
    "box-513x257": ("periodic Rayleigh-Benard box (velvor2dperiodicrbconv.f), 513x257 nodes "
                    "(nr=512 across, np=256 periodic), width 1, Ra=1e5, Pr=1, dt=1e-6",
                    dict(geom=0, nr=512, np=256, width=1.0, dt=1e-6, Ra=1e5, Pr=1.0), 20261203, 1e-3),
    "box-97x337": ("periodic Rayleigh-Benard box (velvor2dperiodicrbconv.f) at the size and physics "
                   "of the reference's own input deck input.s: nr=96, np=336, Ra=1e7, Pr=0.71, "
                   "dt=1e-7, the hard-coded width 3 (np is not a power of two: the direct solve's "
                   "mixed-radix FFT, 336 = 21 x 16)",
                   dict(geom=0, nr=96, np=336, width=3.0, dt=1e-7, Ra=1e7, Pr=0.71), 20261204, 1e-3),
}
DEFAULT_WORKLOAD = "tank-512x256x256"


def check_status(name, info):
    """A run whose solves missed the tolerance (or met a non-finite
    right-hand side) is not a measurement: stop with its status."""
    if info.status != 0:
        sys.exit(f"bench.py: {name}: Poisson status {info.status} (relres v {info.relres_v:.3g}, "
                 f"u {info.relres_u:.3g}); no line is reported")


def gpu_leg(name, steps, warmup, nu, direct_cols=0):
    from codesthesis_dir_amd import BoxConfig, PeriodicBox
    from codesthesis_dir_amd.ic import box_temperature
    desc, kw, seed, amp = WORKLOADS[name]
    cfg = BoxConfig(rtol=1e-10, max_cycles=60, nu1=nu, nu2=nu, direct_cols=direct_cols, **kw)
    box = PeriodicBox(cfg, device=0)
    box.put("th", box_temperature(cfg.nr, cfg.np, cfg.r0, cfg.rmax, amp, seed))
    box.start()
    if warmup:
        box.step(warmup)
    t0 = time.perf_counter()
    info = box.step(steps)          # returns after the stream is idle
    t1 = time.perf_counter()
    wall_ms = (t1 - t0) * 1e3
    check_status(name, info)
    ev_ms = box.last_step_ms()
    # second pass with per-kernel events for the roofline
    box.profile(True)
    box.step(steps)
    prof = box.profile_read()
    box.profile(False)
    nu = float(box.nusselt(info.trans1))
    box.close()
    return cfg, info, wall_ms, ev_ms, prof, nu


MG_KERNELS = ("zebra", "restrict", "prolong", "norm", "direct")


def vcycle_leg(name, steps, warmup, nu=1):
    """The same workload with the Poisson solves by the V-cycle: zebra
    sweeps on the levels of more than 128 columns, the direct level at 128
    (the default is the direct solve of the whole plane up to np = 8192)."""
    from codesthesis_dir_amd import BoxConfig, PeriodicBox
    from codesthesis_dir_amd.ic import box_temperature
    desc, kw, seed, amp = WORKLOADS[name]
    cfg = BoxConfig(rtol=1e-10, max_cycles=60, direct_cols=min(128, kw["np"] // 2), nu1=nu, nu2=nu, **kw)
    box = PeriodicBox(cfg, device=0)
    box.put("th", box_temperature(cfg.nr, cfg.np, cfg.r0, cfg.rmax, amp, seed))
    box.start()
    box.step(warmup)
    t0 = time.perf_counter()
    info = box.step(steps)
    el = time.perf_counter() - t0
    box.close()
    check_status(name + " (V-cycle)", info)
    return {"value": steps / el, "unit": "time-steps/s", "direct_cols": cfg.direct_cols,
            "relres_last_step": [info.relres_v, info.relres_u],
            "vcycle": f"V({nu},{nu}) zebra r-line, phi semi-coarsening",
            "poisson_cycles_last_step": [info.cycles_v, info.cycles_u]}


def lines_scan_leg(name, steps, warmup):
    """The same workload with the compact schemes' line solves in their scan
    form (vv_config.lines = 1: a wavefront per line, the affine recurrences
    composed across the lanes; equal to the default's results to rounding,
    not bit for bit), timed as the default; the compact kernel's device time
    per launch from a second, profiled pass, for both forms."""
    from codesthesis_dir_amd import BoxConfig, PeriodicBox
    from codesthesis_dir_amd.ic import box_temperature
    desc, kw, seed, amp = WORKLOADS[name]
    out = {}
    for lines in (1, 0):
        cfg = BoxConfig(rtol=1e-10, max_cycles=60, lines=lines, **kw)
        box = PeriodicBox(cfg, device=0)
        box.put("th", box_temperature(cfg.nr, cfg.np, cfg.r0, cfg.rmax, amp, seed))
        box.start()
        box.step(warmup)
        t0 = time.perf_counter()
        info = box.step(steps)
        el = time.perf_counter() - t0
        check_status(name + f" (lines={lines})", info)
        box.profile(True)
        box.step(steps)
        p = box.profile_read()["compact"]
        box.profile(False)
        box.close()
        out[lines] = (steps / el, 1e3 * p["ms"] / max(p["launches"], 1))
    return {"value": out[1][0], "unit": "time-steps/s", "lines": "scan (vv_config.lines = 1)",
            "compact_us_per_launch": out[1][1], "reference_lines_same_box": {"value": out[0][0],
            "compact_us_per_launch": out[0][1]}}


def vcycle_large_leg(steps=6, nu=1):
    """V-cycle bandwidth on a large plane: periodic box, 2049 x 4097 nodes
    (nr=2048, np=4096), V(nu,nu) with five smoothed levels and the direct
    level at 128 columns; per-kernel HIP events over `steps` time steps.
    The same plane's time step with the default direct solve is measured
    beside it, and the V-cycle in mixed precision (FP64 defect correction
    around an FP32 cycle): its time, its cycles and residual history, and
    its fields against the FP64 V-cycle run after the same steps."""
    import numpy as np
    from codesthesis_dir_amd import BoxConfig, PeriodicBox
    from codesthesis_dir_amd.ic import box_temperature
    base = dict(nr=2048, np=4096, width=2.0, dt=1e-7, Ra=1e5, Pr=1.0, rtol=1e-10, max_cycles=60)
    th0 = box_temperature(base["nr"], base["np"], 0.5, 1.0, 1e-3, 20261206)

    def vrun(prec):
        cfg = BoxConfig(nu1=nu, nu2=nu, direct_cols=128, precision=prec, **base)
        box = PeriodicBox(cfg, device=0)
        box.put("th", th0)
        box.start()
        box.step(2)
        t0 = time.perf_counter()
        box.step(steps)
        sps = steps / (time.perf_counter() - t0)
        hist = [h.tolist() for h in box.solve_history()]
        box.profile(True)
        info = box.step(steps)
        prof = box.profile_read()
        box.profile(False)
        fields = {f: box.get(f) for f in ("C", "th", "uel", "vel")}
        box.close()
        check_status("box-2049x4097 V-cycle" + (" (mixed)" if prec else ""), info)
        return sps, hist, prof, info, fields

    vstep, vhist, prof, info, f64 = vrun(0)
    mstep, mhist, mprof, minfo, f32 = vrun(1)
    err = {f: float(np.linalg.norm(f32[f] - f64[f]) / np.linalg.norm(f64[f])) for f in f64}
    del f64, f32
    dcfg = BoxConfig(**base)
    dbox = PeriodicBox(dcfg, device=0)
    dbox.put("th", th0)
    dbox.start()
    dbox.step(2)
    t0 = time.perf_counter()
    dinfo = dbox.step(steps)
    dstep = steps / (time.perf_counter() - t0)
    dbox.close()
    check_status("box-2049x4097 direct", dinfo)
    vb = sum(prof[k]["bytes"] for k in MG_KERNELS)
    vt = sum(prof[k]["ms"] for k in MG_KERNELS)
    z = prof["zebra"]
    mt = sum(mprof[k]["ms"] for k in MG_KERNELS)
    return {"workload": "box-2049x4097 (nr=2048, np=4096), V(%d,%d), direct level 128 columns" % (nu, nu),
            "vcycle_gbs": vb / (vt * 1e-3) / 1e9 if vt > 0 else None,
            "vcycle_frac": (vb / (vt * 1e-3) / 1e9) / HBM_PEAK_GBS if vt > 0 else None,
            "zebra_gbs": z["bytes"] / (z["ms"] * 1e-3) / 1e9 if z["ms"] > 0 else None,
            "zebra_avg_launch_us": 1e3 * z["ms"] / max(z["launches"], 1),
            "vcycle_ms_per_step": vt / steps,
            "poisson_cycles_last_step": [info.cycles_v, info.cycles_u],
            "residual_history_last_step": vhist,
            "vcycle_steps_per_s": vstep,
            "mixed_precision": {
                "scheme": "FP64 defect correction (residual, u += e, direct 128-column level, stop test) around "
                          "an FP32 V-cycle (zebra sweeps, restriction, prolongation)",
                "steps_per_s": mstep, "vcycle_ms_per_step": mt / steps,
                "zebra_ms_per_step": mprof["zebra"]["ms"] / steps,
                "fp64_zebra_ms_per_step": z["ms"] / steps,
                "poisson_cycles_last_step": [minfo.cycles_v, minfo.cycles_u],
                "relres_last_step": [minfo.relres_v, minfo.relres_u],
                "residual_history_last_step": mhist,
                "field_rel_l2_vs_fp64_run": err,
                "speedup_vs_fp64_vcycle": mstep / vstep},
            "direct_steps_per_s": dstep,
            "direct_cycles_last_step": [dinfo.cycles_v, dinfo.cycles_u]}


# 3D rotating annulus (rot3dmg.f): the reference program's largest grid
# (PARAMETER kr = kz = 33, kp = 160), 129x129x257 and 257x257x513 (the
# size of a 512x256x256 tank)
ROT3D = {
    "rot3d-33x33x161": (32, 32, 160),
    "rot3d-129x129x257": (128, 128, 256),
    "rot3d-257x257x513": (256, 256, 512),
}
ROT3D_STEPS = {"rot3d-33x33x161": 100, "rot3d-129x129x257": 20, "rot3d-257x257x513": 5}
ROT3D_PHYS = dict(dt=1e-6, Ra=1e7, Re=1000.0, Ptl=7.0, Gr=0.0, r0=0.5, rmax=1.0, zmax=0.5)


def rot3d_state(nr, nz, np_, seed=20261204):
    """Conduction profile T = 1 - ln(r)/ln(r0) (rot3dmg.f:120, rmax = 1) plus
    sum_{m=1..8} a_m cos(m phi + phi_m), a_m ~ U(0, 1e-3), on the interior;
    walls T = 0 (inner) and 1 (outer); fluid at rest."""
    import numpy as np
    P = ROT3D_PHYS
    r0, rmax = P["r0"], P["rmax"]
    dr = (rmax - r0) / nr
    i = np.arange(nr + 2)
    r = np.where(i == 0, r0, np.where(i == nr + 1, rmax, r0 + (i - 0.5) * dr))
    phi = 2.0 * np.pi * np.arange(np_ + 1) / np_
    rng = np.random.default_rng(seed)
    pert = np.zeros(np_ + 1)
    for m in range(1, 9):
        pert += 1e-3 * rng.uniform() * np.cos(m * phi + rng.uniform(0.0, 2.0 * np.pi))
    th = np.broadcast_to(1.0 - np.log(r) / np.log(r0), (np_ + 1, nz + 2, nr + 2)).copy()
    th[:, :, 1:nr + 1] += pert[:, None, None]
    th[:, :, 0], th[:, :, nr + 1] = 0.0, 1.0
    th[0] = th[np_]
    return {"th": th, "thp": th.copy()}


# rot3dmghor.f (stationary horizontal annulus: gravity in the r-theta plane)
# at 257x257x513, the size of a 512x256x256 tank; Gr Ptl = 1e9, Pr = 5
HOR3D = ("rot3dhor-257x257x513", (256, 256, 512), 5)
HOR3D_PHYS = dict(dt=1e-7, Ra=1e9, Re=0.0, Ptl=5.0, Gr=2e8, r0=0.5, rmax=1.0, zmax=0.5, model=1)


def hor3d_state(nr, nz, np_, seed=20261206):
    """rot3dmghor.f's conduction profile T = ln(r)/ln(r0) (rot3dmghor.f:133:
    1 on the inner, 0 on the outer cylinder) plus seeded U(-1e-4, 1e-4) on
    the interior; fluid at rest."""
    import numpy as np
    P = HOR3D_PHYS
    r0, rmax = P["r0"], P["rmax"]
    dr = (rmax - r0) / nr
    i = np.arange(nr + 2)
    r = np.where(i == 0, r0, np.where(i == nr + 1, rmax, r0 + (i - 0.5) * dr))
    th = np.broadcast_to(np.log(r) / np.log(r0), (np_ + 1, nz + 2, nr + 2)).copy()
    rng = np.random.default_rng(seed)
    th[:, :, 1:nr + 1] += rng.uniform(-1e-4, 1e-4, (np_ + 1, nz + 2, nr))
    th[:, :, 0], th[:, :, nr + 1] = 1.0, 0.0
    th[0] = th[np_]
    return {"th": th, "thp": th.copy()}


def rot3d_routines(nr, nz, np_, reps=10):
    """Per-routine device time at one grid: each routine called `reps` times
    back to back (each call synchronises), with its algorithmic bytes (the
    fields it reads and writes) or, for the solves, the transform flops
    (four dense transforms: 2 np^2 m each along phi, 2 nz^2 nr np each
    along z)."""
    from codesthesis_dir_amd.velvor3 import Annulus3D, Annulus3DConfig
    P = ROT3D_PHYS
    run = Annulus3D(Annulus3DConfig(nr=nr, nz=nz, np=np_, **P))
    for k, v in rot3d_state(nr, nz, np_).items():
        run.put(k, v)
    run.step(2)
    N = (nr + 2) * (nz + 2) * (np_ + 1)
    words = {"Avorticity": 9, "Bvorticity": 9, "Cvorticity": 9, "temperature": 6, "rhsw": 3, "rhsv": 3, "rhsu": 4}
    out = {}
    for name, w in words.items():
        f = getattr(run, name)
        f()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        out[name] = {"ms": ms, "GB_s": 8.0 * w * N / (ms * 1e-3) / 1e9}
    m = nr * nz
    for which in (1, 2, 3):
        f = getattr(run, f"poisson{which}")
        f()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        ms = 1e3 * (time.perf_counter() - t0) / reps
        # the dense phi and whole z products' flops: an equivalent rate where the solver
        # uses FFTs or splits the z products by parity
        flops = 2 * (2.0 * np_ * np_ * m) + 2 * (2.0 * nz * nz * nr * np_)
        out[f"poisson{which}"] = {"ms": ms, "transform_TFLOP_s": flops / (ms * 1e-3) / 1e12}
    run.close()
    return out


def rot3d_leg(cpu=True, warmup=3, only=None):
    """The 3D step (vorticity A/B/C, temperature, heat, three exact Poisson
    solves, wall update) on the GPU at both grids, and the reference build
    itself (its multigrid Poisson to the shipped absolute 1e-5) on one core
    at its largest grid when oracle/_ref/librot3d.so is present."""
    import numpy as np
    from codesthesis_dir_amd.velvor3 import Annulus3D, Annulus3DConfig
    P = ROT3D_PHYS
    out = {"physics": "rot3dmg.f, Ra=1e7, Re=1000 (Ta=1e6), Pr=7, Gr=0, dt=1e-6, r0=0.5, rmax=1, zmax=0.5",
           "data": "synthetic: conduction + 8 seeded azimuthal modes (amplitude 1e-3), fluid at rest"}
    for name, (nr, nz, np_) in ROT3D.items():
        if only and name != only:
            continue
        run = Annulus3D(Annulus3DConfig(nr=nr, nz=nz, np=np_, **P))
        for k, v in rot3d_state(nr, nz, np_).items():
            run.put(k, v)
        run.step(warmup)
        n = ROT3D_STEPS[name]
        t0 = time.perf_counter()
        t1 = run.step(n)
        el = time.perf_counter() - t0
        run.close()
        out[name] = {"steps_per_s": n / el, "ms_per_step": 1e3 * el / n, "steps": n, "trans1_last": t1}
        if nr == 128 and not only:
            out[name]["routines"] = rot3d_routines(nr, nz, np_)
    if only:
        return out
    name, (nr, nz, np_), n = HOR3D
    run = Annulus3D(Annulus3DConfig(nr=nr, nz=nz, np=np_, **HOR3D_PHYS))
    for k, v in hor3d_state(nr, nz, np_).items():
        run.put(k, v)
    run.step(warmup)
    t0 = time.perf_counter()
    run.step(n)
    el = time.perf_counter() - t0
    t1, t2 = run.fluxes()
    run.close()
    out[name] = {"steps_per_s": n / el, "ms_per_step": 1e3 * el / n, "steps": n, "trans1_last": t1,
                 "trans2_last": t2,
                 "physics": "rot3dmghor.f (stationary horizontal annulus), Gr=2e8, Pr=5 (Gr Pr = 1e9), dt=1e-7",
                 "data": "synthetic: conduction ln(r)/ln(r0) + seeded U(-1e-4, 1e-4), fluid at rest"}
    lib = os.path.join(ROOT, "oracle", "_ref", "librot3d.so")
    if cpu and os.path.exists(lib):
        nr, nz, np_ = ROT3D["rot3d-33x33x161"]
        with tempfile.TemporaryDirectory() as d:
            st = rot3d_state(nr, nz, np_)
            np.savez(os.path.join(d, "s.npz"), nr=nr, nz=nz, np=np_, **P, f_th=st["th"], f_thp=st["thp"])
            res = os.path.join(d, "o.json")
            subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "refdrv3.py"), os.path.join(d, "s.npz"),
                            res, "8", "10"], stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True)
            r = json.load(open(res))
        out["cpu_reference_33x33x161"] = {
            "value": r["steps"] / r["seconds"], "unit": "time-steps/s", "cores": 1, "kind": "reference",
            "sample": f"{r['steps']} steps of rot3dmg.f compiled as-is (flang -O2), its {r['nlev']}-level "
                      "multigrid to the shipped absolute residual 1e-5, from the same start"}
        g = out["rot3d-33x33x161"]["steps_per_s"]
        out["gpu_over_cpu_reference_33x33x161"] = g / out["cpu_reference_33x33x161"]["value"]
    return out


# sect360v2mg.f (the rotating annulus with a radial barrier) at the
# reference's PARAMETER grid (kr = kz = 33, kp = 161), 129x129x257 and
# 257x257x513, with rot3dmg.f's rotating-annulus physics
SECTOR = {
    "sector-33x33x161": (32, 32, 160),
    "sector-129x129x257": (128, 128, 256),
    "sector-257x257x513": (256, 256, 512),
}
SECTOR_STEPS = {"sector-33x33x161": 100, "sector-129x129x257": 20, "sector-257x257x513": 5}


def sector_state(nr, nz, np_, r0=0.5, rmax=1.0, amp=1e-3, nval=2):
    """sect360v2mg.f's initial temperature (sect360v2mg.f:94-123): the
    conduction profile 1 - ln r / ln r0 (rmax = 1), a mode nval of amplitude
    amp, and the cold line th(nr/4, j, np/2) - 0.2, j = 1..nz; fluid at rest."""
    import numpy as np
    dr = (rmax - r0) / nr
    r = r0 + (np.arange(nr + 2) - 0.5) * dr
    r[0], r[-1] = r0, rmax
    phi = 8.0 * np.arctan(1.0) * np.arange(np_ + 2) / np_
    th = (1.0 - np.log(r) / np.log(r0))[None, None, :] + \
        amp * ((rmax - r) * (r - r0))[None, None, :] * np.cos(nval * phi)[:, None, None]
    th = np.broadcast_to(th, (np_ + 2, nz + 2, nr + 2)).copy()
    th[np_ // 2, 1:nz + 1, nr // 4] -= 0.2
    return {"th": th, "thp": th.copy()}


def sector_leg(warmup=3, only=None, reps=10):
    """The barrier-sector step (include/vvsector.h: A/B/C and temperature in
    one launch, heat, three exact Poisson solves, wall update) at the three
    grids; at 129x129x257 also each routine's time, each call synchronised.
    sect360v2mg.f does not compile here (its include files are absent), so
    there is no reference timing."""
    from codesthesis_dir_amd.sector import Sector, SectorConfig
    P = {k: v for k, v in ROT3D_PHYS.items()}
    out = {"physics": "sect360v2mg.f (radial barrier), Ra=1e7, Re=1000 (Ta=1e6), Pr=7, Gr=0, dt=1e-6, r0=0.5, "
                      "rmax=1, zmax=0.5",
           "data": "synthetic: the reference's start (conduction, mode 2 of amplitude 1e-3, the cold line at "
                   "nr/4, np/2), fluid at rest"}
    for name, (nr, nz, np_) in SECTOR.items():
        if only and name != only:
            continue
        run = Sector(SectorConfig(nr=nr, nz=nz, np=np_, **P))
        for k, v in sector_state(nr, nz, np_).items():
            run.put(k, v)
        run.step(warmup)
        n = SECTOR_STEPS[name]
        t0 = time.perf_counter()
        t1 = run.step(n)
        el = time.perf_counter() - t0
        out[name] = {"steps_per_s": n / el, "ms_per_step": 1e3 * el / n, "steps": n, "trans1_last": t1}
        if nr == 128 and not only:
            N = (nr + 2) * (nz + 2) * (np_ + 2)
            calls = {"vorticity_abc": (lambda: run.vorticity(0), 27), "temperature": (run.temperature, 6),
                     "rhsw": (lambda: run.rhs(1), 3), "rhsv": (lambda: run.rhs(2), 3),
                     "rhsu": (lambda: run.rhs(3), 4), "update": (run.update, 8)}
            rt = {}
            for key, (f, w) in calls.items():
                f()
                t0 = time.perf_counter()
                for _ in range(reps):
                    f()
                ms = 1e3 * (time.perf_counter() - t0) / reps
                rt[key] = {"ms": ms, "GB_s": 8.0 * w * N / (ms * 1e-3) / 1e9}
            for which in (1, 2, 3):
                run.poisson(which)
                t0 = time.perf_counter()
                for _ in range(reps):
                    run.poisson(which)
                ms = 1e3 * (time.perf_counter() - t0) / reps
                k1 = np_ - 1 if which == 2 else np_
                # the whole z products' flops (the phi transforms are FFTs at a power-of-two
                # np): an equivalent rate where the solver splits them by parity (half the flops)
                flops = 2 * (2.0 * nz * nz * nr * k1)
                rt[f"poisson{which}"] = {"ms": ms, "z_transform_TFLOP_s": flops / (ms * 1e-3) / 1e12}
            out[name]["routines"] = rt
        run.close()
    return out


def sweep_requests(n=64, seed=20261205, steps=2000):
    """The served mixed-size sweep: n requests arriving as a Poisson stream
    (mean 20/s); grid sides from {65, 129, 257, 513, 1025} x half height
    (np = side-1, nr = np/2); Ra log-uniform in [1e4, 1e8], Pr uniform in
    [0.7, 7]; Ta = 0 (periodic box, rb.f) or, with probability 1/2,
    log-uniform in [1e3, 1e6] (rotating-annulus plane, Re = sqrt(Ta)); T
    linear in r plus seeded U(-1e-3, 1e-3).

    dt = min(3e-4/sqrt(Ra), 0.05 dr^2), dr = 0.5/nr.  The explicit step has
    a diffusive bound besides the advective one: with the oracle, the
    annulus at nr=128 (Ra=1.1e4, Re=128) grows without bound at dt = 0.18
    dr^2 and stays bounded over 2000 steps at 0.09 dr^2; box runs at nr=512
    diverged at 0.21 dr^2 and above."""
    import numpy as np
    from codesthesis_dir_amd import BoxConfig
    from codesthesis_dir_amd.ic import box_temperature
    from codesthesis_dir_amd.velvor import AnnulusConfig
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    for q in range(n):
        t += rng.exponential(1.0 / 20.0)
        np_ = int(rng.choice([64, 128, 256, 512, 1024]))
        nr = np_ // 2
        Ra = 10.0 ** rng.uniform(4.0, 8.0)
        Pr = rng.uniform(0.7, 7.0)
        Ta = 0.0 if rng.uniform() < 0.5 else 10.0 ** rng.uniform(3.0, 6.0)
        dt = min(3e-4 / Ra ** 0.5, 0.05 * (0.5 / nr) ** 2)
        kw = dict(nr=nr, np=np_, dt=dt, Ra=Ra, Pr=Pr, rtol=1e-10, max_cycles=60, nu1=1, nu2=1,
                  stats=0)
        cfg = AnnulusConfig(Re=Ta ** 0.5, **kw) if Ta > 0 else BoxConfig(width=1.0, **kw)
        th0 = box_temperature(nr, np_, cfg.r0, cfg.rmax, 1e-3, seed + q)
        out.append((t, cfg, th0, steps))
    return out


def serve_leg(n=64, steps=2000, slots=16, baseline=True):
    """The served sweep through the request loop (include/vvserve.h): the
    requests arrive on their schedule and are admitted by slot and device
    footprint; then the same requests one at a time (one slot, all queued
    at once), the plain drop-in's throughput."""
    import numpy as np
    from codesthesis_dir_amd.serve import Server
    reqs = sweep_requests(n, steps=steps)
    res = {}
    with Server(slots=slots, mem_budget=64e9) as srv:
        t_start = time.perf_counter()
        for t_arr, cfg, th0, st in reqs:
            d = t_arr - (time.perf_counter() - t_start)
            if d > 0:
                time.sleep(d)
            srv.submit(cfg, st, th0)
        got = [srv.wait() for _ in reqs]
        wall = time.perf_counter() - t_start
        stats = srv.stats()
    lat = np.array([g["t_done"] - g["t_submit"] for g in got])
    bad = sum(1 for g in got if g["status"] < 0)
    res["failures"] = [{"id": g["id"], "status": g["status"], "error": g["error"], "nr": reqs[g["id"]][1].nr,
                        "np": reqs[g["id"]][1].np, "geom": reqs[g["id"]][1].geom, "Ra": reqs[g["id"]][1].Ra,
                        "Re": reqs[g["id"]][1].Re, "Pr": reqs[g["id"]][1].Pr}
                       for g in got if g["status"] < 0]
    res["served"] = {"slots": slots, "solves_per_s": n / wall, "wall_s": wall, "p50_s": float(np.percentile(lat, 50)),
                     "p99_s": float(np.percentile(lat, 99)), "peak_running": stats["peak_running"],
                     "peak_mem_GB": stats["peak_mem"] / 1e9, "failed": bad,
                     "not_converged": sum(1 for g in got if g["status"] == 1),
                     "runs_created": stats["created"], "runs_reused": stats["reused"]}
    if not baseline:
        return res
    with Server(slots=1, mem_budget=64e9) as srv:
        t_start = time.perf_counter()
        for _, cfg, th0, st in reqs:
            srv.submit(cfg, st, th0)
        got1 = [srv.wait() for _ in reqs]
        wall1 = time.perf_counter() - t_start
    by_side = {}
    for g in got1:
        side = reqs[g["id"]][1].np + 1
        by_side.setdefault(side, []).append(g["t_done"] - g["t_admit"])
    res["one_at_a_time"] = {"solves_per_s": n / wall1, "wall_s": wall1,
                            "failed": sum(1 for g in got1 if g["status"] < 0),
                            "service_s_by_side": {str(k): {"runs": len(v), "mean_s": float(np.mean(v))}
                                                  for k, v in sorted(by_side.items())}}
    res["served_over_one_at_a_time"] = res["served"]["solves_per_s"] / res["one_at_a_time"]["solves_per_s"]
    res["workload"] = (f"{n} requests (served mixed-size sweep), {steps} steps each at Poisson rtol 1e-10, arrivals "
                       "Poisson 20/s, sides {65..1025} x half height, box or annulus plane, FP64")
    return res



# the 3D horizontal-convection tank, 512 x 256 x 256 intervals
# (np = 512 along the tank, nr = 256 across, nz = 256 vertical: aspect 4:1:1)
TANK_WORKLOADS = {
    "tank-512x256x256": (
        "3D horizontal-convection tank (rot3dmg.f's step in Cartesian form, compact derivatives along "
        "all three axes, velvor2dhorconvtank.f's boundaries: linear surface temperature, insulated walls, "
        "stress-free side walls), 512 x 256 x 256 intervals (4:1:1), Ra = Gr Pr = 1e9, Pr = 5, "
        "Ta = 1e8 (Re = 1e4, rotation about the vertical), dt = 1e-7; the three Poisson solves exact "
        "(cosine/sine modes along the tank, z eigenbasis, banded solves across)",
        dict(nr=256, nz=256, np=512, dt=1e-7, Re=1e4, Ptl=5.0, Gr=2e8, lx=1.0, lz=1.0, ly=4.0), 20261206, 1e-4),
}
FP64_MFMA_PEAK_TFS = 78.6


def tank_state(cfg, seed, amp):
    """T = 0.5 + U(-amp, amp) everywhere (seeded), fluid at rest; the first
    update sets the walls."""
    import numpy as np
    rng = np.random.default_rng(seed)
    th = 0.5 + amp * (2.0 * rng.uniform(size=(cfg.np + 2, cfg.nz + 2, cfg.nr + 2)) - 1.0)
    return th


def tank_leg(name, steps, warmup):
    from codesthesis_dir_amd.tank import FLOP_KERNELS, Tank, TankConfig
    desc, kw, seed, amp = TANK_WORKLOADS[name]
    cfg = TankConfig(**kw)
    run = Tank(cfg)
    th = tank_state(cfg, seed, amp)
    run.put("th", th)
    run.put("thp", th)
    del th
    if warmup:
        run.step(warmup)
    t0 = time.perf_counter()
    h = run.step(steps)             # returns after the stream is idle
    wall_ms = (time.perf_counter() - t0) * 1e3
    run.profile(True)
    run.step(steps)
    prof = run.profile_read()
    run.profile(False)
    ke = run.energy()
    run.close()
    for k, v in prof.items():
        v["flop_kernel"] = k in FLOP_KERNELS
    return cfg, wall_ms, prof, h, ke


# the meridional (r,z) plane with swirl on tanh-clustered
# nodes, 513 x 257 nodes, its whole 5000-step run timed
MERID = ("merid-513x257",
         "rotating-annulus meridional (r,z) plane with swirl (velocity-vorticity form of rotann2.f: "
         "azimuthal vorticity, swirl, temperature, Coriolis and centrifugal coupling terms), 513 x 257 "
         "nodes tanh-clustered toward all walls (beta 1.5), radius ratio 0.5, aspect 1, Ra = 1e8, "
         "Ta = 1e5 (Re = sqrt(Ta)), Pr = 1, dt = 2e-8, differentially heated cylinders (T = 0 inside, 1 "
         "outside), insulated lids; fourth-order compact advection derivatives on the uneven nodes, "
         "three-point Fornberg weights elsewhere; both velocity solves exact every step",
         dict(nr=512, nz=256, dt=2e-8, Ra=1e8, Re=1e5 ** 0.5, Pr=1.0, r0=0.5, rmax=1.0, H=0.5, beta=1.5),
         20261203, 1e-4, 5000)


def merid_leg(cpu=True, warmup=20, prof_steps=200, graphs=True):
    """The whole 5000-step run of the meridional workload (timed from the
    seeded start after `warmup` steps of a separate run), then
    `prof_steps` steps with HIP events around every launch."""
    from codesthesis_dir_amd.ic import merid_temperature
    from codesthesis_dir_amd.merid import FLOP_KERNELS, Merid, MeridConfig
    name, desc, kw, seed, amp, steps = MERID
    cfg = MeridConfig(**kw)

    def fresh():
        run = Merid(cfg)
        run.graphs(graphs)
        th = merid_temperature(run.r, cfg.nz, amp, seed)
        for f in ("th", "thp", "thn"):
            run.put(f, th)
        return run

    run = fresh()
    run.step(warmup)                # graph capture and first launches
    run.close()
    run = fresh()
    t0 = time.perf_counter()
    run.step(steps)                 # returns after the stream is idle
    el = time.perf_counter() - t0
    relres = [run.residual(1), run.residual(2)]
    if not all(x <= 1e-10 for x in relres):
        sys.exit(f"bench.py: {name}: velocity solve residuals {relres} above 1e-10; no line is reported")
    heat = run.heat().tolist()
    ke = run.energy()
    run.profile(True)
    run.step(prof_steps)
    prof = run.profile_read()
    run.profile(False)
    run.close()
    for k, v in prof.items():
        v["flop_kernel"] = k in FLOP_KERNELS
    out = {"value": steps / el, "unit": "time-steps/s", "steps": steps, "run_seconds": el,
           "ms_per_step": 1e3 * el / steps,
           "config": {"workload": name, "description": desc, "nr": cfg.nr, "nz": cfg.nz,
                      "poisson": "exact separable solves every step (z eigenbasis, cyclic reduction along r)"},
           "data": "synthetic: T linear in r plus seeded U(-1e-4, 1e-4), fluid at rest in the rotating frame",
           "roofline": tank_roofline(prof, name),
           "kernel_ms_per_step": {k: v["ms"] / prof_steps for k, v in prof.items() if v["launches"]},
           "graph_vs_launches_ms_per_step": [1e3 * el / steps,
                                             sum(v["ms"] for v in prof.values()) / prof_steps],
           "poisson_relres_last_step": relres, "heat_last_step": heat, "kinetic_energy": ke}
    if cpu:
        out["cpu_baseline"] = merid_cpu_leg()
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def merid_cpu_leg(steps=100):
    """`steps` steps of the meridional workload on this host's CPU, one
    thread: the oracle's C restatement of the explicit part and its numpy
    separable solves, in a process of its own."""
    name, desc, kw, seed, amp, _ = MERID
    c = dict(kw, seed=seed, steps=steps)
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "o.json")
        subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "oracle_merid.py"), json.dumps(c), out],
                       check=True, stdout=subprocess.DEVNULL)
        r = json.load(open(out))
    return {"value": r["steps"] / r["seconds"], "unit": "time-steps/s", "cores": 1, "kind": "port",
            "sample": f"{r['steps']} whole steps of the same workload from its seeded start on 1 thread: the "
                      f"oracle's C explicit part and its numpy exact solves ({r['seconds']:.1f} s)"}


def tank_cpu_leg(name):
    """One step of the same workload on this host's CPU, on one thread: the
    oracle's explicit routines (C) and one of its three exact separable
    solves (scipy DCT/DST and numpy eigenbases), in a process of its own."""
    import numpy as np
    desc, kw, seed, amp = TANK_WORKLOADS[name]
    from codesthesis_dir_amd.tank import TankConfig
    cfg = TankConfig(**kw)
    th = tank_state(cfg, seed, amp)
    with tempfile.TemporaryDirectory() as d:
        np.savez(os.path.join(d, "s.npz"), nr=cfg.nr, nz=cfg.nz, np=cfg.np, dt=cfg.dt, Re=cfg.Re, Ptl=cfg.Ptl,
                 Gr=cfg.Gr, lx=cfg.lx, lz=cfg.lz, ly=cfg.ly, f_th=th, f_thp=th)
        del th
        out = os.path.join(d, "o.json")
        subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "oracle_tank.py"), os.path.join(d, "s.npz"),
                        out], check=True, stdout=subprocess.DEVNULL)
        r = json.load(open(out))
    return {"value": 1.0 / r["step_s"], "unit": "time-steps/s", "cores": 1, "kind": "port",
            "sample": f"one step of the workload from its seeded start on 1 thread, timed in two parts: the "
                      f"oracle's C restatement of all the explicit routines ({r['explicit_s']:.1f} s) and one "
                      f"of the three equal-sized exact Poisson solves (scipy DCT-II/DST-I along the tank, "
                      f"numpy eigenbases across; {r['solve_s']:.1f} s); step = explicit + 3 solves"}


def cpu_leg(name, budget_s=15.0):
    """Time the reference CPU path on this host for a bounded sample."""
    desc, kw, seed, amp = WORKLOADS[name]
    nr, np_ = kw["nr"], kw["np"]
    geom = kw.get("geom", 0)
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "rb_omp")
    # the reference's static arrays hold nr <= 289, np <= 1609 (rb.f:54); it
    # runs its hard-coded width 3 (rb.f:132), the same work per step
    if geom == 0 and os.path.exists(ref_bin) and nr <= 289 and np_ <= 1609:
        nts = 200
        with tempfile.TemporaryDirectory() as d:
            deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", str(kw["dt"]), str(nr),
                              str(np_), str(nts), "0.5 1.0 1.0", str(kw["Ra"]), "0.0", "0.0",
                              str(kw["Pr"]), "0.0", "0.0", "1", "no", "n"]) + "\n"
            env = dict(os.environ, OMP_NUM_THREADS="2")
            t0 = time.perf_counter()
            subprocess.run([ref_bin], input=deck, text=True, cwd=d, env=env,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, check=True)
            el = time.perf_counter() - t0
        return {"value": nts / el, "unit": "time-steps/s", "cores": 2, "kind": "reference",
                "sample": f"{nts} steps of the reference executable velvor2dperiodicrbconv.f "
                          f"(flang -O2, 2 OpenMP threads) at {nr}x{np_}"}
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    if geom == 1:
        run = O.Run(nr, np_, kw["dt"], kw["Ra"], kw["Pr"], geom=1, r0=0.5, Re=kw["Re"])
        prog = "velvor2dnewsch3.4.f (static arrays hold nr <= 419)"
    else:
        run = O.Run(nr, np_, kw["dt"], kw["Ra"], kw["Pr"], width=kw["width"])
        prog = "velvor2dperiodicrbconv.f (static arrays hold nr <= 289)"
    run.seed(r0=0.5, amp=amp, seed=seed)
    n = 0
    t0 = time.perf_counter()
    while True:
        st, tr, ncl = run.step(mode=0)
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s or n >= 2000:
            break
    return {"value": n / el, "unit": "time-steps/s", "cores": 1, "kind": "port",
            "sample": f"{n} steps of the same workload from its seeded start: the oracle's C "
                      f"restatement (bit-exact against the flang build) of {prog}, its "
                      f"line-GS FMG Poisson with the shipped absolute stop, 1 thread"}


def copy_calibration(mib):
    """One plain device-to-device copy of `mib` MiB (reads and writes exactly
    that many bytes): the reference dispatch against which the FETCH_SIZE and
    WRITE_SIZE counters are scaled."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def read_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from its counter summary in profiles/
    (a line 'traffic_bytes_per_launch: X'), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def leg_2d(a, name):
    """The 2D headline form of workload `name` (the line of a 2D run)."""
    cfg, info, wall_ms, ev_ms, prof, nu = gpu_leg(name, a.steps, a.warmup, a.nu, a.direct_cols)
    dom = max(prof, key=lambda k: prof[k]["ms"])
    d = prof[dom]
    avg_ms = d["ms"] / max(d["launches"], 1)
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9 if d["ms"] > 0 else 0.0
    vcyc = {k: prof[k] for k in MG_KERNELS}
    vc_ms = sum(v["ms"] for v in vcyc.values())
    vc_bytes = sum(v["bytes"] for v in vcyc.values())
    return {
        "value": a.steps / (wall_ms * 1e-3),
        "unit": "time-steps/s",
        "steps": a.steps,
        "ms_per_step": wall_ms / a.steps,
        "config": {"workload": name, "description": WORKLOADS[name][0],
                   "nr": cfg.nr, "np": cfg.np, "poisson_rtol": cfg.rtol,
                   "poisson": ("direct solve of the whole plane (real FFT in phi, one PCR line solve "
                               "per mode; the default up to np = 8192), checked against 1e-10 each "
                               "step") if cfg.direct_cols in (0, cfg.np)
                              else f"V({cfg.nu1},{cfg.nu2}) zebra r-line, phi semi-coarsening, direct level "
                                   f"at {cfg.direct_cols} columns"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(dom, name),
                     "avg_launch_us": avg_ms * 1e3,
                     "bytes_per_launch": d["bytes"] / max(d["launches"], 1)},
        "vcycle_gbs": vc_bytes / (vc_ms * 1e-3) / 1e9 if vc_ms > 0 else None,
        "kernel_ms_per_step": {k: v["ms"] / a.steps for k, v in prof.items() if v["launches"]},
        "gpu_event_ms_per_step": ev_ms / a.steps,
        "poisson_status": info.status,
        "poisson_relres_last_step": [info.relres_v, info.relres_u],
        "poisson_cycles_last_step": [info.cycles_v, info.cycles_u],
        "nusselt_last_step": nu,
    }


def tank_roofline(prof, name):
    dom = max(prof, key=lambda k: prof[k]["ms"])
    d = prof[dom]
    n = max(d["launches"], 1)
    if d["flop_kernel"]:
        achieved = d["flops"] / (d["ms"] * 1e-3) / 1e12
        return {"bound": "mfma", "kernel": dom, "achieved": achieved, "peak": FP64_MFMA_PEAK_TFS,
                "unit": "TFLOP/s", "frac": achieved / FP64_MFMA_PEAK_TFS, "traffic": read_traffic(dom, name),
                "avg_launch_us": 1e3 * d["ms"] / n, "flops_per_launch": d["flops"] / n,
                "dtype": "f64 (v_mfma_f64_16x16x4f64)"}
    achieved = d["bytes"] / (d["ms"] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(dom, name),
            "avg_launch_us": 1e3 * d["ms"] / n, "bytes_per_launch": d["bytes"] / n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS) + sorted(TANK_WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nu", type=int, default=1, help="zebra sweeps before/after the coarse grid")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the 2D, V-cycle, 3D annulus and served-sweep measurements")
    ap.add_argument("--serve-only", action="store_true",
                    help="measure only the served mixed-size sweep and print its JSON object")
    ap.add_argument("--slots", type=int, default=16, help="request-loop slots of the served sweep")
    ap.add_argument("--direct-cols", type=int, default=0,
                    help="vv_config.direct_cols of a 2D leg (0: the library default)")
    ap.add_argument("--rot3d-grid", default=None, choices=sorted(ROT3D),
                    help="with --rot3d-only: time this grid alone (for profiling)")
    ap.add_argument("--no-baseline", action="store_true",
                    help="with --serve-only: skip the one-at-a-time pass")
    ap.add_argument("--merid-only", action="store_true",
                    help="measure only the meridional plane's 5000-step run and print its JSON object")
    ap.add_argument("--merid-launches", action="store_true",
                    help="with --merid-only: launch each step's kernels one by one instead of replaying graphs")
    ap.add_argument("--vcycle-only", action="store_true",
                    help="measure only the large V-cycle (FP64, mixed precision, direct) and print its JSON object")
    ap.add_argument("--rot3d-only", action="store_true",
                    help="measure only the 3D annulus step (rot3dmg.f) and print its JSON object")
    ap.add_argument("--sector-only", action="store_true",
                    help="measure only the barrier-sector step (sect360v2mg.f) and print its JSON object")
    ap.add_argument("--sector-grid", default=None, choices=sorted(SECTOR),
                    help="with --sector-only: time this grid alone (for profiling)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device copy of this size (counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus}: this build runs on one GPU; runs on several GPUs "
              "(independent runs or requests sharded across devices) are not implemented", file=sys.stderr)
        sys.exit(2)
    if a.rot3d_only:
        print(json.dumps(rot3d_leg(cpu=not a.no_cpu and not a.rot3d_grid, only=a.rot3d_grid)))
        return
    if a.sector_only:
        print(json.dumps(sector_leg(only=a.sector_grid)))
        return
    if a.serve_only:
        print(json.dumps(serve_leg(slots=a.slots, baseline=not a.no_baseline)))
        return
    if a.vcycle_only:
        print(json.dumps(vcycle_large_leg()))
        return
    if a.merid_only:
        print(json.dumps(merid_leg(cpu=not a.no_cpu, graphs=not a.merid_launches)))
        return
    if a.copy_mib > 0:
        copy_calibration(a.copy_mib)

    def done(leg):
        print(f"[bench] {leg} done", file=sys.stderr, flush=True)

    line = {"metric": METRIC, "n_gpus": 1, "warmup": a.warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64"}
    if a.workload in TANK_WORKLOADS:
        cfg, wall_ms, prof, heat, ke = tank_leg(a.workload, a.steps, a.warmup)
        line.update({
            "value": a.steps / (wall_ms * 1e-3),
            "unit": "time-steps/s",
            "steps": a.steps,
            "ms_per_step": wall_ms / a.steps,
            "data": "synthetic: uniform temperature 0.5 plus seeded U(-1e-4, 1e-4), fluid at rest",
            "config": {"workload": a.workload, "description": TANK_WORKLOADS[a.workload][0],
                       "nr": cfg.nr, "nz": cfg.nz, "np": cfg.np,
                       "poisson": "exact solves every step (residual at rounding level)"},
            "roofline": tank_roofline(prof, a.workload),
            "kernel_ms_per_step": {k: v["ms"] / a.steps for k, v in prof.items() if v["launches"]},
            "heat_last_step": list(heat),
            "kinetic_energy": ke,
        })
        done("tank step")
        if not a.no_extra:
            line["merid_513x257"] = merid_leg(cpu=not a.no_cpu)
            done("meridional plane (5000 steps)")
            line["annulus_513x257"] = leg_2d(a, "annulus-513x257")
            done("2D annulus plane")
    else:
        line.update(leg_2d(a, a.workload))
        line["data"] = "synthetic: seeded linear-in-r temperature + uniform splitmix64 noise, fluid at rest"
        done("time step")
    if not a.no_extra:
        name2d = a.workload if a.workload in WORKLOADS else "annulus-513x257"
        line["alternatives"] = {"poisson_vcycle": vcycle_leg(name2d, a.steps, a.warmup, a.nu),
                                "lines_scan": lines_scan_leg(name2d, a.steps, a.warmup)}
        done("V-cycle and scan-line alternatives")
        line["vcycle_large"] = vcycle_large_leg()
        done("large V-cycle")
        line["rot3d"] = rot3d_leg(cpu=not a.no_cpu)
        done("3D annulus step")
        line["sector"] = sector_leg()
        done("barrier-sector step")
        line["served_sweep"] = serve_leg(slots=a.slots)
        done("served sweep")
    if not a.no_cpu:
        line["cpu_baseline"] = tank_cpu_leg(a.workload) if a.workload in TANK_WORKLOADS else cpu_leg(a.workload)
        done("CPU baseline")
    print(json.dumps(line))


if __name__ == "__main__":
    main()
