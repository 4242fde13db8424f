"""Serving engine (cgfdm_server_*) on the GPU, through the C ABI.

* Each request's seismograms and PGV are identical (bit for bit) alone and in a mix of
  requests of different sizes, media, sources and lengths, with the batch launched from a
  HIP graph or directly: the per-point arithmetic does not depend on the batch.
* Served results agree with the single-domain path (cgfdm_step) on the same inputs and the
  scales the server chose for the request within the FP16 target (1e-2 relative L2). They are not bit-identical: the two paths are different
  kernel instantiations of the same stage code, the compiler fuses some multiply-adds
  differently in them (FP32 differences of ~1e-7), and FP16 storage turns last-bit differences
  into rounding trajectories that differ at the FP16 error level (~1e-3).
* An FP32 request served in a mix matches the FP64 oracle to the FP32 target (1e-5 rel L2).
* Admission by memory: with an arena that holds two requests at a time every request still
  completes and no more than two are resident; a request larger than the arena is refused.
* Continuous batching: short requests submitted while long ones run finish first.
* The layered medium built on the device equals the same medium built on the host.
"""
import time

import numpy as np
import pytest

import models
import outputs
from models import rel_l2

pytestmark = pytest.mark.gpu


def _requests(cg, prec, seed=11, n=6, grid=None):
    rng = np.random.default_rng(seed)
    reqs = []
    for q in range(n):
        NX, NY, NZ = (int(x) for x in rng.integers([36, 34, 28], [64, 60, 44]))
        if grid is not None:
            NX, NY, NZ = grid
        NT = int(rng.integers(30, 70))
        lay = (np.array([0.0, 1200.0, 2500.0][: 2 + q % 2], np.float32),
               np.array([1800.0, 2600.0, 3200.0][: 2 + q % 2], np.float32), None, None)
        vp = (1.73 * lay[1]).astype(np.float32)
        rho = (1740.0 * (vp / 1000.0) ** 0.25).astype(np.float32)
        lay = (lay[0], lay[1], vp, rho)
        d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=NT, nPML=6, precision=prec,
                         use_terrain=q % 3 == 2, depth=NZ * 100.0)
        r = {"desc": d, "layers": lay,
             "stations": models.surface_stations(NX, NY, NZ, n=3)}
        if d.use_terrain:
            r["terrain"] = models.gauss_hill(NX, NY, 100.0, h0=250.0, width=900.0)
        if q % 2 == 0:
            r["point"] = (NX // 2, NY // 2, NZ - 12, 2.5)
        else:
            mu = float(rho[-1]) * float(lay[1][-1]) ** 2
            import workloads as wl
            tab = wl.double_couple_table(1e15, 30.0 * q, 60.0, 10.0 * q, 2.0, NT + 2, 0.008, mu)
            r["table"] = (np.array([[NX // 2 - 2, NY // 2 + 1, NZ - 10]], np.int32), tab, 0.008)
        reqs.append(r)
    return reqs


def _serve(cg, reqs, prec, **kw):
    out = {}
    with cg.Server(precision=prec, **kw) as s:
        ids = {s.submit(r): q for q, r in enumerate(reqs)}
        while len(out) < len(reqs):
            got = s.wait(timeout_ms=60000)
            assert got, f"no request completed within 60 s ({s.stats()})"
            for rid in got:
                rc, st, pgv, info = s.result(rid)
                assert rc == 0, info["error"]
                out[ids[rid]] = (st, pgv, info)
        stats = s.stats()
    return out, stats


def _served_scales(cg, r):
    """The scales the server chooses for a 16-bit request without its own (cgfdm_source_scales
    with the largest shear modulus of the layers and of the default medium; a point source:
    cgfdm_point_source_scales with the smallest shear modulus of the layers)."""
    tops, vs, vp, rho = r["layers"]
    mus = np.asarray(rho, np.float64) * np.asarray(vs, np.float64) ** 2
    mu = max(float(mus.max()), 2670.0 * 3464.0 ** 2)
    if "point" in r:
        return cg.point_source_scales(r["desc"].DH, r["point"][3], float(mus.min()), mu)
    xyz, tab, dt = r["table"]
    return cg.source_scales(r["desc"].DH, xyz, tab, dt, mu)


def _alone(cg, r):
    """The request through cgfdm_step, with the scales the server chooses for it."""
    import copy
    d = copy.copy(r["desc"])
    if d.precision != cg.FP32 and d.Cv == 0 and d.Cs == 0:
        d.Cv, d.Cs = _served_scales(cg, r)
    g = cg.Domain(d)
    g.set_model_layered(r.get("terrain"), *r["layers"])
    if "point" in r:
        g.set_source_point(*r["point"])
    else:
        g.set_source_table(*r["table"])
    g.set_stations(r["stations"])
    g.step(d.NT)
    st, pgv = g.stations(), g.pgv()
    g.close()
    return st, pgv


@pytest.mark.parametrize("use_graph", [1, 0])
def test_same_alone_and_in_the_mix(gpu, use_graph):
    cg = gpu
    reqs = _requests(cg, cg.FP16)
    mix, stats = _serve(cg, reqs, cg.FP16, max_domains=8, steps_per_launch=8, use_graph=use_graph)
    assert stats["completed"] == len(reqs) and stats["failed"] == 0
    assert stats["peak_resident"] > 1 and stats["graph"] == use_graph
    for q, r in enumerate(reqs):
        alone, _ = _serve(cg, [r], cg.FP16, max_domains=1, steps_per_launch=4, use_graph=use_graph)
        assert np.array_equal(mix[q][0], alone[0][0]), f"request {q}: stations differ alone/mix"
        assert np.array_equal(mix[q][1], alone[0][1]), f"request {q}: PGV differs alone/mix"
        st1, pgv1 = _alone(cg, r)
        assert np.isfinite(st1).all() and np.abs(st1).max() > 0
        assert rel_l2(mix[q][0], st1) <= 1e-2, f"request {q}: stations vs cgfdm_step"
        assert rel_l2(mix[q][1], pgv1) <= 1e-2, f"request {q}: PGV vs cgfdm_step"


def test_served_fp32_matches_oracle(gpu, oracle_lib):
    cg, po = gpu, oracle_lib
    reqs = _requests(cg, cg.FP32, seed=5, n=4)
    mix, _ = _serve(cg, reqs, cg.FP32, max_domains=4, steps_per_launch=8, use_graph=1)
    for q in (0, 1):
        r = reqs[q]
        d = r["desc"]
        o = po.Oracle(d, po.F64)
        NX, NY, NZ = d.NX, d.NY, d.NZ
        z = models.terrain_z(r["terrain"], NZ, d.depth) if d.use_terrain else models.flat_z(NX, NY, NZ, 100.0)
        h = r["terrain"][3:-3, 3:-3] if d.use_terrain else np.zeros((NY, NX))
        vs, vp, rho = outputs.layered_medium(z, np.broadcast_to(h, z.shape), *(np.asarray(a, np.float64) for a in r["layers"]), from_km=False)
        o.set_model(r.get("terrain"), vs, vp, rho)
        if "point" in r:
            o.set_source_point(*r["point"])
        else:
            o.set_source_table(*r["table"])
        o.set_stations(r["stations"])
        o.step(d.NT)
        assert rel_l2(mix[q][0], o.stations()) <= 1e-5, q
        assert rel_l2(mix[q][1], o.pgv()) <= 1e-5, q


def test_admission_by_memory(gpu):
    cg = gpu
    reqs = _requests(cg, cg.FP16, seed=3, n=5, grid=(48, 44, 36))
    first, _ = _serve(cg, reqs, cg.FP16, max_domains=8)
    biggest = max(v[2]["device_bytes"] for v in first.values())
    budget = int(2.2 * biggest)
    again, stats = _serve(cg, reqs, cg.FP16, max_domains=8, hbm_budget=budget)
    assert stats["completed"] == len(reqs)
    assert stats["peak_resident"] <= 2 and stats["arena_peak"] <= budget
    for q in range(len(reqs)):
        assert np.array_equal(first[q][0], again[q][0])
    with cg.Server(precision=cg.FP16, max_domains=2, hbm_budget=budget) as s:
        big = dict(reqs[0])
        big["desc"] = cg.make_desc(300, 300, 150, NT=10, nPML=6, precision=cg.FP16)
        with pytest.raises(cg.CgfdmError, match="ENOMEM"):
            s.submit(big)


def test_short_requests_overtake_long_ones(gpu):
    cg = gpu
    base = _requests(cg, cg.FP16, seed=9, n=4)
    long_, short = base[:2], base[2:]
    for r in long_:
        r["desc"].NT = 600
        if "table" in r:
            xyz, tab, dt = r["table"]
            r["table"] = (xyz, np.concatenate([tab, np.zeros((600, 1, 6), np.float32)]), dt)
    for r in short:
        r["desc"].NT = 16
    done = {}
    with cg.Server(precision=cg.FP16, max_domains=8, steps_per_launch=4) as s:
        ids = {s.submit(r): ("long", q) for q, r in enumerate(long_)}
        t0 = time.time()
        while s.stats()["resident"] < 2:
            assert time.time() - t0 < 30, "long requests never became resident"
            time.sleep(0.001)
        ids.update({s.submit(r): ("short", q) for q, r in enumerate(short)})
        while len(done) < 4:
            got = s.wait(timeout_ms=60000)
            assert got, f"no request completed within 60 s ({s.stats()})"
            for rid in got:
                rc, st, pgv, info = s.result(rid)
                assert rc == 0, info["error"]
                done[ids[rid]] = info
    # the short ones were admitted while the long ones ran, and finished before them
    for q in range(2):
        assert done[("short", q)]["t_admit"] < min(done[("long", p)]["t_done"] for p in range(2))
        assert done[("short", q)]["t_done"] < min(done[("long", p)]["t_done"] for p in range(2))


def test_layered_model_on_device_equals_host(gpu):
    cg = gpu
    NX, NY, NZ = 40, 36, 30
    lay = (np.array([0.0, 900.0, 1800.0], np.float32), np.array([1700.0, 2400.0, 3300.0], np.float32),
           np.array([3000.0, 4200.0, 5700.0], np.float32), np.array([2100.0, 2400.0, 2600.0], np.float32))
    for use_terrain in (0, 1):
        d = cg.make_desc(NX, NY, NZ, DH=100.0, NT=4, nPML=6, use_terrain=use_terrain, depth=NZ * 100.0)
        terrain = models.gauss_hill(NX, NY, 100.0, h0=400.0, width=900.0) if use_terrain else None
        z = models.terrain_z(terrain, NZ, d.depth) if use_terrain else models.flat_z(NX, NY, NZ, 100.0)
        h = terrain[3:-3, 3:-3] if use_terrain else np.zeros((NY, NX))
        vs, vp, rho = outputs.layered_medium(z, np.broadcast_to(h, z.shape), *(np.asarray(a, np.float64) for a in lay), from_km=False)
        a = cg.Domain(d)
        a.set_model_layered(terrain, *lay)
        b = cg.Domain(d)
        b.set_model(terrain, vs, vp, rho)
        assert np.array_equal(a.cjm(), b.cjm()), use_terrain
        a.close()
        b.close()


@pytest.mark.parametrize("least_work_first", [False, True])
def test_admission_order(gpu, least_work_first):
    """One slot, held by a long request while two more wait: the free slot takes the earlier
    submitted one in arrival order, the one with less work (NX*NY*NZ*NT) when asked."""
    cg = gpu
    first, a, b = _requests(cg, cg.FP16, seed=5, n=3)  # point, table, point sources
    first["desc"].NT = 2000  # holds the slot while both others are submitted
    a["desc"].NT = 16  # within its table's length
    b["desc"].NT = 64

    def work(r):
        d = r["desc"]
        return d.NX * d.NY * d.NZ * d.NT

    big, small = (a, b) if work(a) > work(b) else (b, a)
    infos = {}
    with cg.Server(precision=cg.FP16, max_domains=1, steps_per_launch=4, least_work_first=least_work_first) as s:
        ids = {s.submit(first): "first"}
        t0 = time.time()
        while s.stats()["resident"] < 1:
            assert time.time() - t0 < 30, "the first request never became resident"
            time.sleep(0.001)
        ids[s.submit(big)] = "big"
        ids[s.submit(small)] = "small"
        st = s.stats()
        assert st["resident"] == 1 and st["pending"] == 2, st  # both wait for the one slot
        while len(infos) < 3:
            got = s.wait(timeout_ms=60000)
            assert got, f"no request completed within 60 s ({s.stats()})"
            for rid in got:
                rc, st, pgv, info = s.result(rid)
                assert rc == 0, info["error"]
                infos[ids[rid]] = info
    assert infos["first"]["t_done"] <= min(infos["big"]["t_admit"], infos["small"]["t_admit"])
    if least_work_first:
        assert infos["small"]["t_admit"] < infos["big"]["t_admit"]
    else:
        assert infos["big"]["t_admit"] < infos["small"]["t_admit"]


def test_served_non_finite_result_is_reported(gpu):
    """A 16-bit request whose own scales put its field out of range (Cv 1e9: velocities of mm/s
    overflow FP16) retires with CGFDM_ENONFINITE and its outputs still copied; a request served
    beside it is unaffected."""
    cg = gpu
    bad, good = _requests(cg, "fp16", seed=5, n=2, grid=(40, 38, 30))
    bad["desc"].Cv, bad["desc"].Cs = 1e9, 1e9 / 5e9
    with cg.Server(precision="fp16", max_domains=4, rescale_retries=0) as s:
        ids = {s.submit(bad): "bad", s.submit(good): "good"}
        seen = {}
        while len(seen) < 2:
            got = s.wait(timeout_ms=60000)
            assert got, f"no request completed within 60 s ({s.stats()})"
            for rid in got:
                seen[ids[rid]] = s.result(rid)
    rc, st, pgv, info = seen["bad"]
    assert rc == cg.STATUS_CODES["ENONFINITE"] and "non-finite" in info["error"]
    assert not np.isfinite(st).all()
    rc, st, pgv, info = seen["good"]
    assert rc == 0 and np.isfinite(st).all() and np.isfinite(pgv).all()


def test_non_finite_request_is_run_again_with_smaller_scales(gpu):
    """With rescale_retries, the request of the test above (Cv 1e9) is run again from its start
    with Cv and Cs divided by 8 until its results are finite: it completes with status OK, its
    info names the scales of the run that produced the results and how many runs were repeated,
    and the results are bit-identical to the same request served once with those scales."""
    cg = gpu
    bad, good = _requests(cg, "fp16", seed=5, n=2, grid=(40, 38, 30))
    bad["desc"].Cv, bad["desc"].Cs = 1e9, 1e9 / 5e9
    with cg.Server(precision="fp16", max_domains=4, rescale_retries=8) as s:
        ids = {s.submit(bad): "bad", s.submit(good): "good"}
        seen = {}
        while len(seen) < 2:
            got = s.wait(timeout_ms=60000)
            assert got, f"no request completed within 60 s ({s.stats()})"
            for rid in got:
                seen[ids[rid]] = s.result(rid)
        stats = s.stats()
    rc, st, pgv, info = seen["bad"]
    k = info["rescaled"]
    print(f"rescaled {k} times, Cv {info['Cv']:.4g}")
    assert rc == 0 and k >= 1 and stats["rescaled"] == k and stats["completed"] == 2
    assert info["Cv"] == pytest.approx(1e9 / 8 ** k, rel=1e-12)
    assert info["Cs"] == pytest.approx(1e9 / 5e9 / 8 ** k, rel=1e-12)
    assert np.isfinite(st).all() and np.isfinite(pgv).all()
    assert seen["good"][0] == 0 and seen["good"][3]["rescaled"] == 0
    again, _ = _requests(cg, "fp16", seed=5, n=2, grid=(40, 38, 30))
    again["desc"].Cv, again["desc"].Cs = info["Cv"], info["Cs"]
    with cg.Server(precision="fp16", max_domains=4, rescale_retries=0) as s:
        rid = s.submit(again)
        assert s.wait(timeout_ms=60000) == [rid]
        rc2, st2, pgv2, info2 = s.result(rid)
    assert rc2 == 0 and info2["rescaled"] == 0
    assert np.array_equal(st2, st) and np.array_equal(pgv2, pgv)
