"""The compact schemes' line solves in their scan form (vv_config.lines =
VV_LINES_SCAN: one wavefront per line, the reference's elimination factors
with the affine recurrences composed by a scan across the lanes) against
the reference form (bit-exact with the oracle, tests/test_gpu_parity.py):

* one explicit step (both vorticity and temperature, whose derivatives come
  from the six line families) agrees with the reference form to 1e-13
  relative, on box and annulus planes whose lines span 1 to 32 rows per
  lane, np a power of two or not;
* 1000 steps of the 129 x 65 box with the scan form stay within 1e-9
  (relative L2) of the oracle's trajectory with exact solves, the bar of
  the V-cycle and mixed-precision runs.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _run(geom, nr, np_, lines, **kw):
    from codesthesis_dir_amd import AnnulusConfig, BoxConfig, PeriodicBox
    print(f"[gpu] lines={lines} geom={geom} nr={nr} np={np_}", flush=True)
    if geom:
        return PeriodicBox(AnnulusConfig(nr=nr, np=np_, Re=100.0, lines=lines, **kw), device=0)
    return PeriodicBox(BoxConfig(nr=nr, np=np_, width=1.0, lines=lines, **kw), device=0)


@pytest.mark.parametrize("geom", [0, 1])
@pytest.mark.parametrize("nr,np_", [(8, 16), (64, 128), (512, 256), (96, 336), (1000, 64), (2048, 40),
                                    (40, 2048)])
def test_scan_lines_match_reference_lines(geom, nr, np_):
    rng = np.random.default_rng(nr + 7 * np_ + geom)
    ld = nr + 2
    state = {n: rng.standard_normal((np_ + 1, ld)) for n in ("C", "Cp", "th", "thp", "uel", "vel")}
    for a in state.values():
        a[0] = a[np_]
    out = {}
    for lines in (0, 1):
        run = _run(geom, nr, np_, lines, dt=1e-5, Ra=1e5, Pr=1.0)
        for n, a in state.items():
            run.put(n, a)
        run.cvorticity()
        run.temperature()
        out[lines] = {n: run.get(n) for n in ("Cn", "thn", "C", "th")}
        run.close()
    for n in ("Cn", "thn", "C", "th"):
        ref, got = out[0][n], out[1][n]
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        assert err <= 1e-13, (n, err)


def test_scan_lines_trajectory_1000_steps():
    """The 129 x 65 box (Ra = 1e5, Pr = 1, dt = 1e-5) with the scan form:
    every field within 1e-9 of the oracle after 1000 steps."""
    nr, np_, width = 64, 128, 1.0
    run = O.Run(nr, np_, 1e-5, 1e5, 1.0, width=width)
    run.seed()
    box = _run(0, nr, np_, 1, dt=1e-5, Ra=1e5, Pr=1.0, rtol=1e-10, max_cycles=60)
    box.put("th", run["th"])
    box.start()
    info = box.step(1000)
    assert info.status == 0
    for _ in range(1000):
        st, tr, ncl = run.step(mode=1)
        assert st == 0
    for name in ("th", "C", "uel", "vel", "Cp", "thp"):
        g, o = box.get(name), run[name]
        err = np.linalg.norm(g - o) / np.linalg.norm(o)
        print(f"{name}: rel L2 {err:.3e}")
        assert err <= 1e-9, (name, err)
    box.close()
