"""GPU parity of every kernel instantiation and workload configuration that produces numbers.

  tile shapes      the bulk kernel forced to each tile shape (64x8, 64x4, 64x16, 128x8) and
                   z-chunk length (automatic, 4 planes) on a grid whose bulk box spans three
                   128-wide x tiles and several y tiles and chunks, FP32 (<= 1e-5 against the FP64
                   and FP32 oracles) and FP16 (<= 1e-2 against the FP64 oracle)
  terrain, 16-bit  the Gaussian-hill workload at reduced size: hill, 3-layer medium, PML, free
                   surface, strike-slip double couple through the moment-rate table; FP16 and BF16
                   storage against the FP64 oracle (FP16 <= 1e-2 relative L2 on the seismograms
                   and PGV; BF16 <= 5e-2)
  strong source    the reduced Gaussian-hill workload at M0 = 1e17 N m, which overflows FP16 at
                   the reference's fixed scales: FP16 storage with scales chosen from the source
                   (cgfdm_source_scales) against the FP64 oracle (<= 1e-2)
  Gaussian hill    the full-size Gaussian-hill workload (401x401x201, 2000 steps, M0 = 1e17 N m):
                   FP16 storage with scales from the source finite and within 1e-2 of FP32
  finite fault     the kinematic finite-fault workload (workloads.finite_fault_workload): at 500 m
                   spacing (241 x 241 x 81, 800 steps) FP32 (<= 1e-5) and FP16 with scales from the
                   source (<= 1e-2) against the FP64 oracle's output in
                   tests/golden/config5_reduced_fp64.npz; at full size (1201 x 1201 x 401) the first
                   600 steps in FP16 finite and within 1e-2 of FP32
  half space       the full-size half-space workload (201x201x101, 1000 steps, 21x21 stations)
                   against the FP64 oracle's output stored in tests/golden/config1_fp64.npz
                   (<= 1e-5)
  throughput       the full-size throughput workload (1001x1001x401: random terrain, perturbed
                   layered medium, double couple), FP32 and FP16 storage for 400 steps: finite,
                   FP16 seismograms within 1e-2 of FP32
"""
import os

import numpy as np
import pytest

import models
import workloads as wl
from models import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SCALES16 = {"Cv": 5e3, "Cs": 1e-6, "Crho": 1e6}


# ---------------------------------------------------------------- bulk tile shapes
@pytest.mark.parametrize("kc", [0, 4])
@pytest.mark.parametrize("cfg", [0, 1, 2, 3])
def test_bulk_tile_shapes_fp32(gpu, oracle_lib, monkeypatch, cfg, kc):
    cg, po = gpu, oracle_lib
    # nPML 6: the bulk box is 258 x 18 x 15 (x: two full 128-wide tiles and a ragged third)
    NX, NY, NZ = 270, 30, 24
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=8, nPML=6)
    med = models.homogeneous(NX, NY, NZ)
    monkeypatch.setenv("CGFDM_BULK_CFG", str(cfg))
    monkeypatch.setenv("CGFDM_BULK_KC", str(kc))
    g = cg.Domain(d)
    g.set_model(None, *med)
    o64 = po.Oracle(d, po.F64)
    o64.set_model(None, *med)
    o32 = po.Oracle(d, po.F32)
    o32.set_model(None, *med)
    w0 = models.smooth_state(NX, NY, NZ, seed=17).astype(np.float32)
    g.set_state(w0)
    o64.set_state(w0.astype(np.float64))
    o32.set_state(w0.astype(np.float64))
    for n in range(3):
        g.step(1)
        o64.step(1)
        o32.step(1)
        wg = g.state()
        assert rel_l2(wg, o64.state()) <= 1e-5, (n, rel_l2(wg, o64.state()))
        assert rel_l2(wg, o32.state()) <= 1e-5, (n, rel_l2(wg, o32.state()))
    bulk_pts, _, tile, kchunk = g.partition()
    assert (tile, kchunk) == (cfg, kc)
    assert bulk_pts == 258 * 18 * 15
    g.close()


@pytest.mark.parametrize("cfg", [0, 1, 2, 3])
def test_bulk_tile_shapes_fp16(gpu, oracle_lib, monkeypatch, cfg):
    cg, po = gpu, oracle_lib
    NX, NY, NZ = 270, 30, 24
    d16 = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=8, nPML=6, precision="fp16", **SCALES16)
    d64 = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=8, nPML=6, **SCALES16)
    med = models.homogeneous(NX, NY, NZ)
    monkeypatch.setenv("CGFDM_BULK_CFG", str(cfg))
    monkeypatch.setenv("CGFDM_BULK_KC", "0")
    g = cg.Domain(d16)
    g.set_model(None, *med)
    o = po.Oracle(d64, po.F64)
    o.set_model(None, *med)
    w0 = po.round_f16(models.smooth_state(NX, NY, NZ, seed=19, Cv=5e3, Cs=1e-6).astype(np.float32))
    g.set_state(w0)
    o.set_state(w0.astype(np.float64))
    g.step(3)
    o.step(3)
    e = rel_l2(g.state(), o.state())
    print(f"tile cfg {cfg}: FP16 state after 3 steps, rel L2 vs FP64 oracle {e:.2e}")
    assert e <= 1e-2
    assert g.partition()[2] == cfg
    g.close()


# ---------------------------------------------------------------- Gaussian hill, 16-bit storage
def _hill_case(cg, prec, scales, M0=1e15):
    """Reduced Gaussian-hill workload: 72 x 68 x 44 at DH 100 m, hill 1500 m * (6/40) high and
    6000 m * (72/400) wide, layers at 3 and 10 km scaled by the same 44/201 depth ratio."""
    NX, NY, NZ, DH, DT, NT = 72, 68, 44, 100.0, 0.006, 360
    depth = NZ * DH
    terrain = wl.gauss_hill(NX, NY, DH, h0=225.0, width=1080.0)
    z = models.terrain_z(terrain, NZ, depth)
    med = models.layered_on_depth(z - terrain[3:-3, 3:-3][None], [0.0, 0.66, 2.19], [2000.0, 3000.0, 3500.0])
    d = cg.make_desc(NX, NY, NZ, DH=DH, DT=DT, NT=NT, nPML=10, precision=prec, use_terrain=1, depth=depth,
                     **scales)
    # strike-slip double couple (strike 0, dip 90, rake 0), 1.1 km below the hill top
    X, Y, Z = NX // 2, NY // 2, NZ - 1 - 11
    vs, vp, rho = med
    mu = float(rho[Z, Y, X]) * float(vs[Z, Y, X]) ** 2
    table = wl.double_couple_table(M0, 0.0, 90.0, 0.0, 2.0, NT + 2, DT, mu)
    src = np.array([[X, Y, Z]], np.int32)
    st = wl.surface_stations(NX, NY, NZ, n=7, margin=12)
    return d, terrain, med, src, table, DT, st, NT


@pytest.fixture(scope="module")
def hill_fp64(oracle_lib):
    """The FP64 oracle's stations and PGV for the reduced hill workload (run once per module)."""
    import cgfdm
    po = oracle_lib
    d64, terrain, med, src, table, dt, st, NT = _hill_case(cgfdm, "fp32", SCALES16)
    o = po.Oracle(d64, po.F64)
    o.set_model(terrain, *med)
    o.set_source_table(src, table, dt)
    o.set_stations(st)
    o.step(NT)
    return o.stations(), o.pgv()


@pytest.mark.parametrize("prec,tol", [("fp16", 1e-2), ("bf16", 5e-2)])
def test_hill_layered_table_source_low_precision_vs_fp64(gpu, hill_fp64, prec, tol):
    cg = gpu
    d, terrain, med, src, table, dt, st, NT = _hill_case(cg, prec, SCALES16)
    g = cg.Domain(d)
    g.set_model(terrain, *med)
    g.set_source_table(src, table, dt)
    g.set_stations(st)
    g.step(NT)
    sg = g.stations()
    so, po_ = hill_fp64
    assert np.isfinite(sg).all() and np.abs(so[:, :, :3]).max() > 0
    ev = rel_l2(sg[:, :, :3], so[:, :, :3])
    et = rel_l2(sg[:, :, 3:], so[:, :, 3:])
    ep = rel_l2(g.pgv(), po_)
    peak = np.abs(sg[:, :, :3]).max() / np.abs(so[:, :, :3]).max() - 1.0
    print(f"{prec} hill: velocity seismograms {ev:.2e}, stress {et:.2e}, PGV {ep:.2e}, peak {peak:+.2e}")
    assert ev <= tol and ep <= tol
    g.close()


@pytest.mark.parametrize("prec,tol", [("fp16", 1e-2), ("bf16", 5e-2)])
def test_hill_strong_source_scales_from_source_vs_fp64(gpu, oracle_lib, prec, tol):
    """M0 = 1e17 N m, a hundred times the moment the reference's fixed scales are used with here:
    FP16 and BF16 storage with the scales chosen from the source match the FP64 oracle."""
    cg, po = gpu, oracle_lib
    M0 = 1e17
    d64, terrain, med, src, table, dt, st, NT = _hill_case(cg, "fp32", SCALES16, M0=M0)
    o = po.Oracle(d64, po.F64)
    o.set_model(terrain, *med)
    o.set_source_table(src, table, dt)
    o.set_stations(st)
    o.step(NT)
    so = o.stations()
    print(f"M0 {M0:g}: oracle peak surface |v| {np.abs(so[:, :, :3]).max():.3g} m/s")
    vs, vp, rho = med
    Cv, Cs = cg.source_scales(100.0, src, table, dt, float((rho * vs * vs).max()))
    d, *_ = _hill_case(cg, prec, {"Cv": Cv, "Cs": Cs, "Crho": 1e6}, M0=M0)
    g = cg.Domain(d)
    g.set_model(terrain, *med)
    g.set_source_table(src, table, dt)
    g.set_stations(st)
    g.step(NT)
    sg = g.stations()
    assert np.isfinite(sg).all()
    ev = rel_l2(sg[:, :, :3], so[:, :, :3])
    ep = rel_l2(g.pgv(), o.pgv())
    print(f"{prec} hill M0 1e17, Cv {Cv:.4g} Cs {Cs:.4g}: seismograms {ev:.2e}, PGV {ep:.2e} vs FP64")
    assert ev <= tol and ep <= tol
    g.close()
    o.close()


# ---------------------------------------------------------------- Gaussian hill, full size
@pytest.mark.timeout(600)
def test_gauss_hill_full_size_fp16_scales_from_source_vs_fp32(gpu):
    """The Gaussian-hill workload at full size: 401 x 401 x 201, DH 100 m, hill 1500 m high and
    6 km wide, layers Vs 2000/3000/3500 m/s below 3 and 10 km, strike-slip double couple
    (strike 0, dip 90, rake 0) of M0 = 1e17 N m 5 km deep, 2 Hz Ricker, 2000 steps of 10 ms."""
    cg = gpu
    NX, NY, NZ, DH, DT, NT = 401, 401, 201, 100.0, 0.01, 2000
    depth = NZ * DH
    terrain = wl.gauss_hill(NX, NY, DH, h0=1500.0, width=6000.0)
    z = models.terrain_z(terrain, NZ, depth)
    med = models.layered_on_depth(z - terrain[3:-3, 3:-3][None], [0.0, 3.0, 10.0], [2000.0, 3000.0, 3500.0])
    vs, vp, rho = med
    X, Y, Z = NX // 2, NY // 2, NZ - 1 - 50
    mu = float(rho[Z, Y, X]) * float(vs[Z, Y, X]) ** 2
    table = wl.double_couple_table(1e17, 0.0, 90.0, 0.0, 2.0, NT + 2, DT, mu)
    src = np.array([[X, Y, Z]], np.int32)
    st = wl.surface_stations(NX, NY, NZ, n=21)
    Cv, Cs = cg.source_scales(DH, src, table, DT, float((rho * vs * vs).max()))
    out = {}
    for prec, sc in (("fp32", {}), ("fp16", {"Cv": Cv, "Cs": Cs, "Crho": 1e6})):
        d = cg.make_desc(NX, NY, NZ, DH=DH, DT=DT, NT=NT, nPML=12, precision=prec, use_terrain=1, depth=depth,
                         check_cfl=1, **sc)
        g = cg.Domain(d)
        g.set_model(terrain, *med)
        g.set_source_table(src, table, DT)
        g.set_stations(st)
        g.step(NT)
        out[prec] = (g.stations(), g.pgv())
        g.close()
        print(f"Gaussian hill {prec}: {NT} steps done", flush=True)
    s32, p32 = out["fp32"]
    s16, p16 = out["fp16"]
    assert np.isfinite(s16).all() and np.isfinite(p16).all()
    e = rel_l2(s16[:, :, :3], s32[:, :, :3])
    ep = rel_l2(p16, p32)
    peak = float(np.abs(s32[:, :, :3]).max())
    print(f"Gaussian hill 401x401x201, M0 1e17 (Cv {Cv:.4g}): peak surface |v| {peak:.3g} m/s, FP16 vs FP32 "
          f"seismograms {e:.2e}, PGV {ep:.2e}")
    assert peak > 0 and e <= 1e-2 and ep <= 1e-2


# ---------------------------------------------------------------- kinematic finite fault
def _fault_table(cg, w, d, tmp_path):
    path = tmp_path / "fault.bin"
    wl.write_rupture(path, w["rupture"])
    return cg.finite_fault_map(path, d, w["terrain"], w["lon0"], w["lat0"])


def _fault_scales(cg, w, xyz, table, tdt):
    vs, vp, rho = w["medium"]
    return cg.source_scales(w["DH"], xyz, table, tdt, float((rho.astype(np.float64) * vs * vs).max()))


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-5), ("fp16", 1e-2)])
def test_finite_fault_reduced_vs_fp64_golden(gpu, tmp_path, prec, tol):
    cg = gpu
    gold = np.load(os.path.join(HERE, "golden", "config5_reduced_fp64.npz"))
    w = wl.finite_fault_workload(DH=500.0)
    NX, NY, NZ = w["grid"]
    NT = int(gold["nt"])
    d32 = cg.make_desc(NX, NY, NZ, DH=w["DH"], DT=w["DT"], NT=NT, nPML=w["nPML"], use_terrain=1, depth=w["depth"])
    xyz, table, tdt = _fault_table(cg, w, d32, tmp_path)
    np.testing.assert_array_equal(xyz, gold["source_xyz"])  # the library maps the file as the oracle
    sc = {}
    if prec != "fp32":
        Cv, Cs = _fault_scales(cg, w, xyz, table, tdt)
        sc = {"Cv": Cv, "Cs": Cs, "Crho": 1e6}
    d = cg.make_desc(NX, NY, NZ, DH=w["DH"], DT=w["DT"], NT=NT, nPML=w["nPML"], precision=prec, use_terrain=1,
                     depth=w["depth"], check_cfl=1, **sc)
    st = w["stations"]
    sel = np.array([r * 20 + c for r in range(0, 20, 2) for c in range(0, 20, 2)])
    np.testing.assert_array_equal(st[sel], gold["station_xyz"])
    g = cg.Domain(d)
    g.set_model(w["terrain"], *w["medium"])
    g.set_source_table(xyz, table, tdt)
    g.set_stations(st)
    g.step(NT)
    s = g.stations()[sel, :, :3]
    e = rel_l2(s, gold["stations"])
    ep = rel_l2(g.pgv(), gold["pgv"])
    print(f"finite fault 241x241x81 {prec} {sc}: seismograms {e:.2e}, PGV {ep:.2e} vs FP64 oracle")
    assert np.abs(gold["stations"]).max() > 0 and np.isfinite(s).all()
    assert e <= tol and ep <= tol
    g.close()


@pytest.mark.timeout(600)
def test_finite_fault_full_size_fp16_vs_fp32(gpu, tmp_path):
    cg = gpu
    w = wl.finite_fault_workload(DH=100.0)
    NX, NY, NZ = w["grid"]
    NT = 600
    d32 = cg.make_desc(NX, NY, NZ, DH=w["DH"], DT=w["DT"], NT=NT, nPML=w["nPML"], use_terrain=1, depth=w["depth"])
    xyz, table, tdt = _fault_table(cg, w, d32, tmp_path)
    assert len(xyz) == 1600
    Cv, Cs = _fault_scales(cg, w, xyz, table, tdt)
    out = {}
    for prec, sc in (("fp32", {}), ("fp16", {"Cv": Cv, "Cs": Cs, "Crho": 1e6})):
        d = cg.make_desc(NX, NY, NZ, DH=w["DH"], DT=w["DT"], NT=NT, nPML=w["nPML"], precision=prec, use_terrain=1,
                         depth=w["depth"], check_cfl=1, **sc)
        g = cg.Domain(d)
        g.set_model(w["terrain"], *w["medium"])
        g.set_source_table(xyz, table, tdt)
        g.set_stations(w["stations"])
        g.step(NT)
        out[prec] = (g.stations(), g.pgv())
        g.close()
        print(f"finite fault 1201x1201x401 {prec}: {NT} steps done", flush=True)
    s32, p32 = out["fp32"]
    s16, p16 = out["fp16"]
    assert np.isfinite(s16).all() and np.isfinite(p16).all()
    e = rel_l2(s16[:, :, :3], s32[:, :, :3])
    ep = rel_l2(p16, p32)
    print(f"finite fault 1201x1201x401, {NT} steps (Cv {Cv:.4g}): peak surface |v| {np.abs(s32[:, :, :3]).max():.3g} m/s, "
          f"FP16 vs FP32 seismograms {e:.2e}, PGV {ep:.2e}")
    assert np.abs(p32).max() > 0
    assert e <= 1e-2 and ep <= 1e-2


# ---------------------------------------------------------------- half space, full size
def test_half_space_full_size_vs_fp64_golden(gpu):
    cg = gpu
    gold = np.load(os.path.join(HERE, "golden", "config1_fp64.npz"))
    NX, NY, NZ, NT = 201, 201, 101, int(gold["nt"])
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=NT, nPML=12)
    st = wl.surface_stations(NX, NY, NZ, n=21)
    sel = np.array([r * 21 + c for r in range(0, 21, 2) for c in range(0, 21, 2)])
    np.testing.assert_array_equal(st[sel], gold["station_xyz"])
    g = cg.Domain(d)
    g.set_model(None, *models.homogeneous(NX, NY, NZ))
    g.set_source_point(100, 100, 80, 2.0)
    g.set_stations(st)
    g.step(NT)
    s = g.stations()[sel, :, :3]
    e = rel_l2(s, gold["stations"])
    ep = rel_l2(g.pgv(), gold["pgv"])
    print(f"half space 201x201x101, {NT} steps: seismograms rel L2 {e:.2e}, PGV {ep:.2e} vs FP64 oracle")
    assert np.abs(gold["stations"]).max() > 0
    assert e <= 1e-5 and ep <= 1e-5
    g.close()


# ---------------------------------------------------------------- throughput workload, full size
@pytest.mark.timeout(600)
def test_throughput_full_size_fp16_vs_fp32(gpu):
    cg = gpu
    NX, NY, NZ, DH, DT, NT = 1001, 1001, 401, 100.0, 0.005, 400
    terrain = wl.smooth_random_terrain(NX, NY, DH, rms=1000.0, corr=5000.0, seed=2024)
    depth = NZ * DH
    vs, vp, rho = wl.build_medium(NX, NY, NZ, DH, terrain=terrain, depth=depth, pert=0.05, corr_m=1000.0, seed=2025)
    sk = NZ - 1 - 50
    mu = float(rho[sk, NY // 2, NX // 2]) * float(vs[sk, NY // 2, NX // 2]) ** 2
    table = wl.double_couple_table(1e16, 0.0, 90.0, 0.0, 2.0, NT + 2, DT, mu)
    src = np.array([[NX // 2, NY // 2, sk]], np.int32)
    st = wl.surface_stations(NX, NY, NZ, n=21)
    out = {}
    for prec in ("fp32", "fp16"):
        d = cg.make_desc(NX, NY, NZ, DH=DH, DT=DT, NT=NT, nPML=12, precision=prec, use_terrain=1, depth=depth)
        g = cg.Domain(d)
        g.set_model(terrain, vs, vp, rho)
        g.set_source_table(src, table, DT)
        g.set_stations(st)
        g.step(NT)
        out[prec] = (g.stations(), g.pgv())
        g.close()
        print(f"throughput {prec}: {NT} steps done", flush=True)
    s32, p32 = out["fp32"]
    s16, p16 = out["fp16"]
    assert np.isfinite(s16).all() and np.isfinite(p16).all() and np.isfinite(s32).all()
    assert np.abs(s32[:, :, :3]).max() > 0
    e = rel_l2(s16[:, :, :3], s32[:, :, :3])
    ep = rel_l2(p16, p32)
    print(f"throughput 1001x1001x401, {NT} steps: FP16 vs FP32 seismograms {e:.2e}, PGV {ep:.2e}")
    assert e <= 1e-2 and ep <= 1e-2
