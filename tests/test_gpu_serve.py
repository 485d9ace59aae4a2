"""The request loop (include/vvserve.h, codesthesis_dir_amd/serve.py) on the
GPU: a run's results are the same alone and served among others, and
admission keeps the admitted footprint within the budget."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests():
    from codesthesis_dir_amd import BoxConfig
    from codesthesis_dir_amd.ic import box_temperature
    from codesthesis_dir_amd.velvor import AnnulusConfig
    out = []
    for q, (nr, np_, geom) in enumerate([(16, 32, 0), (32, 64, 1), (32, 128, 0), (64, 64, 1), (16, 64, 0),
                                         (64, 256, 1)]):
        kw = dict(nr=nr, np=np_, dt=2e-5, Ra=1e5, Pr=1.0, rtol=1e-10, max_cycles=60)
        cfg = AnnulusConfig(Re=100.0, **kw) if geom else BoxConfig(width=1.0, **kw)
        out.append((cfg, box_temperature(nr, np_, cfg.r0, cfg.rmax, 1e-3, 50 + q)))
    return out


def test_results_same_alone_and_in_the_mix():
    """Six runs of different sizes and geometries served four at a time give
    the fields and fluxes of the same runs made alone through the plain
    API, bit for bit."""
    from codesthesis_dir_amd import PeriodicBox
    from codesthesis_dir_amd.serve import Server
    reqs = _requests()
    nsteps = 40
    with Server(slots=4, mem_budget=8e9) as srv:
        ids = [srv.submit(cfg, nsteps, th0, outputs=True) for cfg, th0 in reqs]
        got = {}
        for _ in ids:
            g = srv.wait(timeout=120.0)
            assert g is not None
            got[g["id"]] = g
        st = srv.stats()
    assert st["peak_running"] >= 2
    for rid, (cfg, th0) in zip(ids, reqs):
        g = got[rid]
        assert g["status"] == 0, g["status"]
        box = PeriodicBox(cfg, device=0)
        box.put("th", th0)
        box.start()
        info = box.step(nsteps)
        th, vel = box.get("th"), box.get("vel")
        box.close()
        print(f"[gpu] request {rid}: nr={cfg.nr} np={cfg.np} geom={cfg.geom}", flush=True)
        assert np.array_equal(g["th"], th[:, :cfg.nr + 2])
        assert np.array_equal(g["vel"], vel[:, :cfg.nr + 2])
        assert g["trans1"] == info.trans1 and g["trans2"] == info.trans2


def test_admission_keeps_the_budget():
    """With room for two runs' footprints only two run at once, and a
    request larger than the budget is refused at submission."""
    from codesthesis_dir_amd._lib import VVError
    from codesthesis_dir_amd.serve import Server, footprint
    cfg, th0 = _requests()[2]
    f = footprint(cfg)
    with Server(slots=6, mem_budget=2.5 * f) as srv:
        for _ in range(5):
            srv.submit(cfg, 20, th0)
        for _ in range(5):
            assert srv.wait(timeout=120.0)["status"] == 0
        st = srv.stats()
        assert st["peak_running"] <= 2
        assert st["peak_mem"] <= 2.5 * f
        big, th_big = _requests()[5]
        if footprint(big) > 2.5 * f:
            with pytest.raises(VVError):
                srv.submit(big, 10, th_big)


def test_resident_runs_give_fresh_results():
    """A run kept resident and re-armed for a request of the same grid
    (vv_reset: new physics and initial state, tables kept) gives the
    fields of a run created for it, bit for bit; a re-arm for another grid
    is refused."""
    from codesthesis_dir_amd import PeriodicBox
    from codesthesis_dir_amd._lib import VVError
    from codesthesis_dir_amd.serve import Server
    from dataclasses import replace
    reqs = _requests()
    cfg_a, th_a = reqs[1]                               # annulus 32 x 64
    cfg_b = replace(cfg_a, Ra=3e5, dt=1e-5, Re=40.0)    # same grid, other physics
    th_b = th_a[::-1].copy()
    fresh = {}
    for name, cfg, th0 in (("a", cfg_a, th_a), ("b", cfg_b, th_b)):
        box = PeriodicBox(cfg, device=0)
        box.put("th", th0)
        box.start()
        info = box.step(30)
        fresh[name] = (box.get("th"), box.get("vel"), info.trans1)
        box.close()
    # the same through one resident run: a, b, a
    with Server(slots=1, mem_budget=8e9) as srv:
        for cfg, th0 in ((cfg_a, th_a), (cfg_b, th_b), (cfg_a, th_a)):
            srv.submit(cfg, 30, th0, outputs=True)
        res = [srv.wait(timeout=120.0) for _ in range(3)]
        st = srv.stats()
    assert st["created"] == 1 and st["reused"] == 2 and st["resident"] == 1
    for g, name in zip(res, ("a", "b", "a")):
        th, vel, t1 = fresh[name]
        assert g["status"] == 0
        assert np.array_equal(g["th"], th[:, :cfg_a.nr + 2]), name
        assert np.array_equal(g["vel"], vel[:, :cfg_a.nr + 2]), name
        assert g["trans1"] == t1, name
    # the plain API: a run re-armed after other steps equals a fresh one
    box = PeriodicBox(cfg_b, device=0)
    box.put("th", th_b)
    box.start()
    box.step(17)
    box.reset(cfg_a)
    box.put("th", th_a)
    box.start()
    info = box.step(30)
    assert np.array_equal(box.get("th"), fresh["a"][0]) and info.trans1 == fresh["a"][2]
    with pytest.raises(VVError):
        box.reset(reqs[0][0])                           # another grid
    box.close()
