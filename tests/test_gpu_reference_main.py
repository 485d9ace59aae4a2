"""The reference's own main programs, unmodified, on the GPU.

oracle/Makefile links the main programs of velvor2dperiodicrbconv.f (lines
1-577), velvor2dnewsch3.4.f (1-589), rot3dmg.f (1-333) and rot3dmghor.f
(1-347) against the Fortran drop-ins (libvvdropin, libvvdropin_ann,
libvvdropin3, libvvdropin3h) into oracle/_ref/, next to the reference's own
executables.
Both run the same input deck; their printed heat-transfer series and the
field files they write are compared.  The drop-ins solve the Poisson
problems exactly (2D: to a relative residual of 1e-10), the reference stops
its multigrid at an absolute residual (2e-6 in 2D, rb.f:2086; 1e-5 in 3D,
rot3dmg.f:1562), so the bounds below are the reference's own solver error.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref")

pytestmark = pytest.mark.gpu


def _run(exe, deck, d):
    print(f"[gpu] {os.path.basename(exe)} in {d}", flush=True)
    res = subprocess.run([exe], input=deck, text=True, capture_output=True, cwd=d, timeout=300)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    return res.stdout


def _series(out, ncol):
    """Rows of the per-step print: ncol floats then the step number."""
    rows = []
    for line in out.splitlines():
        t = line.split()
        if len(t) == ncol + 1 and t[-1].isdigit():
            try:
                rows.append([float(x) for x in t[:ncol]])
            except ValueError:
                pass
    return np.array(rows)


def _records(path):
    """Records of a Fortran sequential unformatted file (4-byte markers)."""
    raw = open(path, "rb").read()
    out, p = [], 0
    while p < len(raw):
        n = int(np.frombuffer(raw[p:p + 4], dtype=np.int32)[0])
        out.append(raw[p + 4:p + 4 + n])
        p += n + 8
    return out


def _field(path, rec):
    return np.frombuffer(_records(path)[rec], dtype=np.float64)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "r3_main_gpu")),
                    reason="reference main programs not linked (reference sources were absent at build)")
@pytest.mark.parametrize("grid", [("16", "16", "32"), ("32", "32", "160")])
def test_rot3dmg_main_program_on_the_gpu(grid):
    """rot3dmg.f's main program (its deck, initial condition, time loop and
    output files) through the 3D drop-in against the reference executable;
    32 x 32 x 160 is the largest grid its static arrays hold, where
    np = 160 = 5 * 32 takes the mixed-radix FFT of the phi transforms."""
    nts = 10
    deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "1e-5", *grid, str(nts),
                      "0.5 1.0 0.5", "1e5", "100.0", "0.0", "1.0", "0.0", "0.1", "4", "no", "n"]) + "\n"
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg:
        out_r = _run(os.path.join(REF, "r3_ref"), deck, dr)
        out_g = _run(os.path.join(REF, "r3_main_gpu"), deck, dg)
        hr, hg = _series(out_r, 1), _series(out_g, 1)
        assert hr.shape == hg.shape == (nts, 1), (hr.shape, hg.shape)
        eh = _rel(hg, hr)
        errs = {n: _rel(_field(os.path.join(dg, n), 2), _field(os.path.join(dr, n), 2))
                for n in ("T", "A", "B", "C", "U", "V", "W")}
    print(f"rot3dmg main {'x'.join(grid)}: heat series rel {eh:.3e}; fields {errs}")
    # measured on the MI355X: heat 1.4e-12, T 7.5e-12, A..W <= 6e-6
    assert eh < 1e-9
    assert errs["T"] < 1e-9
    for n in ("A", "B", "C", "U", "V", "W"):
        assert errs[n] < 1e-4, (n, errs[n])


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "r3h_main_gpu")),
                    reason="reference main programs not linked (reference sources were absent at build)")
@pytest.mark.parametrize("grid", [("16", "16", "32"), ("32", "32", "160")])
def test_rot3dmghor_main_program_on_the_gpu(grid):
    """rot3dmghor.f's main program (the stationary horizontal annulus: its
    deck, initial condition, time loop with heat(trans1,trans2) and output
    files) through the drop-in built for it (libvvdropin3h.so) against the
    reference executable, also at 32 x 32 x 160 (np = 5 * 32: the
    mixed-radix FFT).  The reference stops its solves at an absolute
    residual of 1e-6 (rot3dmghor.f:1599)."""
    nts = 10
    deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "1e-5", *grid, str(nts),
                      "0.5 1.0 0.5", "1e5", "100.0", "1e4", "1.0", "0.0", "0.1", "4", "no", "n"]) + "\n"
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg:
        out_r = _run(os.path.join(REF, "r3h_ref"), deck, dr)
        out_g = _run(os.path.join(REF, "r3h_main_gpu"), deck, dg)
        hr, hg = _series(out_r, 2), _series(out_g, 2)
        assert hr.shape == hg.shape == (nts, 2), (hr.shape, hg.shape)
        eh = _rel(hg, hr)
        errs = {n: _rel(_field(os.path.join(dg, n), 2), _field(os.path.join(dr, n), 2))
                for n in ("T", "A", "B", "C", "U", "V", "W")}
    print(f"rot3dmghor main {'x'.join(grid)}: heat series rel {eh:.3e}; fields {errs}")
    assert eh < 1e-9
    assert errs["T"] < 1e-9
    for n in ("A", "B", "C", "U", "V", "W"):
        assert errs[n] < 1e-4, (n, errs[n])


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "r3_main_gpu")),
                    reason="reference main programs not linked (reference sources were absent at build)")
def test_rot3dmg_main_program_resident_and_full_copies():
    """rot3dmg.f's main program at 32 x 32 x 160, 40 steps, through the 3D
    drop-in with its default residency (a step returns A, B, C, th, uel,
    vel, wel) and with full copies (VV_DROPIN_RESIDENT=0): the printed flux
    series and every file the program writes are the same bit for bit; the
    wall time of each run is printed."""
    import time
    nts = 40
    deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "1e-5", "32", "32", "160", str(nts),
                      "0.5 1.0 0.5", "1e5", "100.0", "0.0", "1.0", "0.0", "0.1", "4", "no", "n"]) + "\n"
    exe = os.path.join(REF, "r3_main_gpu")
    out, secs, files = {}, {}, {}
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d0:
        for mode, d in (("1", d1), ("0", d0)):
            print(f"[gpu] r3_main_gpu VV_DROPIN_RESIDENT={mode} in {d}", flush=True)
            t = time.perf_counter()
            res = subprocess.run([exe], input=deck, text=True, capture_output=True, cwd=d, timeout=300,
                                 env=dict(os.environ, VV_DROPIN_RESIDENT=mode))
            secs[mode] = time.perf_counter() - t
            assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
            out[mode] = _series(res.stdout, 1)
            files[mode] = {n: open(os.path.join(d, n), "rb").read() for n in ("T", "A", "B", "C", "U", "V", "W")}
    print(f"rot3dmg main 32x32x160, {nts} steps: resident {secs['1']:.2f} s, full copies {secs['0']:.2f} s")
    assert out["1"].shape == (nts, 1)
    assert np.array_equal(out["1"], out["0"])
    for n, b in files["1"].items():
        assert b == files["0"][n], n


def _plane_main(ref_exe, gpu_exe, deck, nts, formatted=False):
    """Runs a 2D program's executable and its main program on the GPU with
    one deck; returns the heat-series error, the field-file errors and the
    error of the running-mean file zmtmtempvar.dat.  The field files are
    unformatted records, or (formatted=True) one E16.10 value per line."""
    read = (lambda path: np.loadtxt(path)) if formatted else (lambda path: _field(path, 0))
    env_one = dict(os.environ, OMP_NUM_THREADS="1")
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg:
        print(f"[gpu] {os.path.basename(ref_exe)} in {dr}", flush=True)
        res = subprocess.run([ref_exe], input=deck, text=True, capture_output=True, cwd=dr, timeout=300,
                             env=env_one)
        assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
        out_g = _run(gpu_exe, deck, dg)
        hr, hg = _series(res.stdout, 2), _series(out_g, 2)
        assert hr.shape == hg.shape == (nts // 50, 2), (hr.shape, hg.shape)
        errs = {n: _rel(read(os.path.join(dg, n)), read(os.path.join(dr, n))) for n in ("T", "C", "U", "V")}
        mr = np.loadtxt(os.path.join(dr, "zmtmtempvar.dat"))
        mg = np.loadtxt(os.path.join(dg, "zmtmtempvar.dat"))
        return _rel(hg, hr), errs, _rel(mg, mr)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ann_main_gpu")),
                    reason="reference main programs not linked (reference sources were absent at build)")
def test_annulus_main_program_on_the_gpu():
    """velvor2dnewsch3.4.f's main program (the rotating-annulus plane)
    through the annulus build of the 2D drop-in against the reference
    executable."""
    nts = 200   # the program prints its heat fluxes every 50 steps (velvor2dnewsch3.4.f:391)
    deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "1e-5", "32", "64", str(nts), "0.5 1.0 1.0",
                      "1e5", "100.0", "0.0", "1.0", "0.0", "0.0", "1", "no", "n"]) + "\n"
    eh, errs, em = _plane_main(os.path.join(REF, "ann_ref"), os.path.join(REF, "ann_main_gpu"), deck, nts,
                               formatted=True)
    print(f"annulus main: heat series rel {eh:.3e}; fields {errs}; running means {em:.3e}")
    assert eh < 1e-9
    assert errs["T"] < 1e-9
    for n in ("C", "U", "V"):
        assert errs[n] < 1e-4, (n, errs[n])
    assert em < 1e-4


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "rb_main_gpu")),
                    reason="reference main programs not linked (reference sources were absent at build)")
def test_periodic_box_main_program_on_the_gpu():
    """velvor2dperiodicrbconv.f's main program through the 2D drop-in against
    the reference executable (run with one OpenMP thread)."""
    nts = 200   # the program prints its heat fluxes every 50 steps (rb.f:391)
    deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "1e-5", "32", "64", str(nts), "0.5 1.0 1.0",
                      "1e5", "0.0", "0.0", "1.0", "0.0", "0.0", "1", "no", "n"]) + "\n"
    eh, errs, em = _plane_main(os.path.join(REF, "rb_omp"), os.path.join(REF, "rb_main_gpu"), deck, nts)
    print(f"rb main: heat series rel {eh:.3e}; fields {errs}; running means {em:.3e}")
    # measured on the MI355X: heat 1.5e-12, T 4.6e-12, C, U, V <= 1.4e-7
    assert eh < 1e-9
    assert errs["T"] < 1e-9
    for n in ("C", "U", "V"):
        assert errs[n] < 1e-5, (n, errs[n])
    assert em < 1e-5


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "rb_main_gpu")),
                    reason="reference main programs not linked (reference sources were absent at build)")
def test_periodic_box_main_program_on_its_own_deck():
    """velvor2dperiodicrbconv.f's main program on the reference's own input
    deck (input.s: dt 1e-7, nr 96, np 336, Ra 1e7, Pr 0.71, pert 0,
    amp 1, nval 3), 200 steps instead of 50000, output file names of this
    test, through the 2D drop-in against the reference executable.  np = 336
    is not a power of two: the drop-in's direct solve uses the dense
    transforms over the 336 columns."""
    nts = 200
    deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "1.d-7", "96", "336", str(nts),
                      "0.5d0,1.0d0,0.1d0", "1.d7", "0d0", "0.d0", "0.71d0", "0.d0", "1.d0", "3", "no",
                      "n"]) + "\n"
    eh, errs, em = _plane_main(os.path.join(REF, "rb_omp"), os.path.join(REF, "rb_main_gpu"), deck, nts)
    print(f"rb main, input.s deck: heat series rel {eh:.3e}; fields {errs}; running means {em:.3e}")
    assert eh < 1e-9
    assert errs["T"] < 1e-9
    for n in ("C", "U", "V"):
        assert errs[n] < 1e-5, (n, errs[n])
    assert em < 1e-5


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "rb_main_gpu")),
                    reason="reference main programs not linked (reference sources were absent at build)")
def test_periodic_box_main_program_resident_and_full_copies():
    """The input.s deck (96 x 336), 1000 steps, through the drop-in with its
    default residency (a step returns C, th, uel and vel to COMMON) and with
    full copies of every routine's inputs and outputs (VV_DROPIN_RESIDENT=0):
    the printed heat series and every file the program writes are the same
    bit for bit; the wall time of each run is printed."""
    import time
    nts = 1000
    deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "1.d-7", "96", "336", str(nts),
                      "0.5d0,1.0d0,0.1d0", "1.d7", "0d0", "0.d0", "0.71d0", "0.d0", "1.d0", "3", "no",
                      "n"]) + "\n"
    exe = os.path.join(REF, "rb_main_gpu")
    out, secs, files = {}, {}, {}
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d0:
        for mode, d in (("1", d1), ("0", d0)):
            print(f"[gpu] rb_main_gpu VV_DROPIN_RESIDENT={mode} in {d}", flush=True)
            t = time.perf_counter()
            res = subprocess.run([exe], input=deck, text=True, capture_output=True, cwd=d, timeout=300,
                                 env=dict(os.environ, VV_DROPIN_RESIDENT=mode))
            secs[mode] = time.perf_counter() - t
            assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
            out[mode] = _series(res.stdout, 2)
            files[mode] = {n: open(os.path.join(d, n), "rb").read()
                           for n in ("T", "A", "B", "C", "U", "V", "W", "zmtmtempvar.dat")}
    print(f"rb main, input.s deck, {nts} steps: resident {secs['1']:.2f} s, full copies {secs['0']:.2f} s")
    assert out["1"].shape == (nts // 50, 2)
    assert np.array_equal(out["1"], out["0"])
    for n, b in files["1"].items():
        assert b == files["0"][n], n


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "rb_main_gpu_omp")),
                    reason="reference main programs not linked (reference sources were absent at build)")
def test_periodic_box_main_program_with_openmp_sections():
    """The same main program built with OpenMP and run on 2 threads: its
    sections call rhsv/poisson2 and rhsu/poisson3 at once (rb.f:370-380),
    which reach one run of the drop-in from two threads.  The results match
    the reference executable (one thread) as closely as the serial build's."""
    nts = 200
    deck = "\n".join(["A", "B", "C", "U", "V", "W", "T", "no", "1e-5", "32", "64", str(nts), "0.5 1.0 1.0",
                      "1e5", "0.0", "0.0", "1.0", "0.0", "0.0", "1", "no", "n"]) + "\n"
    env2 = dict(os.environ, OMP_NUM_THREADS="2")
    with tempfile.TemporaryDirectory() as dr, tempfile.TemporaryDirectory() as dg:
        res = subprocess.run([os.path.join(REF, "rb_omp")], input=deck, text=True, capture_output=True, cwd=dr,
                             timeout=300, env=dict(os.environ, OMP_NUM_THREADS="1"))
        assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
        print(f"[gpu] rb_main_gpu_omp in {dg}, 2 threads", flush=True)
        out = subprocess.run([os.path.join(REF, "rb_main_gpu_omp")], input=deck, text=True, capture_output=True,
                             cwd=dg, timeout=300, env=env2)
        assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
        hr, hg = _series(res.stdout, 2), _series(out.stdout, 2)
        assert hr.shape == hg.shape == (nts // 50, 2)
        errs = {n: _rel(_field(os.path.join(dg, n), 0), _field(os.path.join(dr, n), 0)) for n in ("T", "C", "U", "V")}
    eh = _rel(hg, hr)
    print(f"rb main, OpenMP sections: heat series rel {eh:.3e}; fields {errs}")
    assert eh < 1e-9
    assert errs["T"] < 1e-9
    for n in ("C", "U", "V"):
        assert errs[n] < 1e-5, (n, errs[n])


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "ann_main_gpu")),
                    reason="reference main programs not linked (reference sources were absent at build)")
def test_annulus_main_program_on_its_own_deck(golden_dir):
    """velvor2dnewsch3.4.f's main program on its shipped deck
    input1e81e5pr1-1.s (nr 240, np 1792, Ra 1e8, Re 1e5, Pr 0.1), 50 steps
    from the program's own start (the deck restarts from field files the
    reference does not ship), through the annulus drop-in, against the
    reference executable's output on the same deck
    (ann_deck_240x1792.npz, written by make_golden_anndeck.py).  np = 1792 =
    7 * 256 takes the mixed-radix FFT of the direct solve."""
    import sys
    sys.path.insert(0, golden_dir)
    import make_golden_anndeck as M
    g = np.load(os.path.join(golden_dir, "ann_deck_240x1792.npz"))
    stride = int(g["stride"])
    with tempfile.TemporaryDirectory() as d:
        out = _run(os.path.join(REF, "ann_main_gpu"), M.DECK, d)
        eh = _rel(_series(out, 2), g["heat"])
        errs = {}
        for f in ("T", "C", "U", "V"):
            v = np.loadtxt(os.path.join(d, f))
            assert v.size == int(g[f"n_{f}"]), f
            errs[f] = float(np.max(np.abs(v[::stride] - g[f"s_{f}"])) / float(g[f"norm_{f}"]))
        em = _rel(np.loadtxt(os.path.join(d, "zmtmtempvar.dat")), g["means"])
    print(f"annulus main, input1e81e5pr1-1.s at 240x1792: heat {eh:.3e}; fields {errs}; running means {em:.3e}")
    assert eh < 1e-9
    assert errs["T"] < 1e-9
    for n in ("C", "U", "V"):
        assert errs[n] < 1e-4, (n, errs[n])
    assert em < 1e-4
