"""GPU parity: the HIP path through the C ABI against the CPU oracle on the same inputs.

Tolerances (relative L2 over the compared array):
  FP32 storage vs FP64 oracle       <= 1e-5   (agreed FP32 target)
  FP32 storage vs FP32 restatement  <= 1e-5
  FP16 storage vs FP64 oracle       <= 1e-2   (agreed FP16 target)
  BF16 storage vs FP64 oracle       <= 5e-2   (8-bit mantissa; reported, no agreed target)
"""
import numpy as np
import pytest

import models
from models import rel_l2

pytestmark = pytest.mark.gpu


def _pair(cg, po, desc, prec_oracle, terrain=None, medium=None):
    d = desc
    vs, vp, rho = medium if medium is not None else models.homogeneous(d.NX, d.NY, d.NZ)
    g = cg.Domain(d)
    g.set_model(terrain, vs, vp, rho)
    o = po.Oracle(d, prec_oracle)
    o.set_model(terrain, vs, vp, rho)
    return g, o


def test_cjm_flat_and_terrain(gpu, oracle_lib):
    cg, po = gpu, oracle_lib
    NX, NY, NZ = 30, 26, 20
    for use_terrain in (0, 1):
        d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=10, nPML=5, use_terrain=use_terrain,
                         depth=NZ * 100.0)
        terrain = models.gauss_hill(NX, NY, 100.0, h0=300.0, width=800.0) if use_terrain else None
        g, o = _pair(cg, po, d, po.F64, terrain)
        a = g.cjm().astype(np.float64)
        b = o.cjm64().astype(np.float32).astype(np.float64)
        # terrain-following grids: xi_y, xi_z, eta_x, eta_z are exactly zero on both sides (the
        # stage tile kernels leave their terms out)
        for c in (1, 2, 3, 5):
            assert not a[c].any() and not b[c].any(), f"terrain={use_terrain}: metric component {c} not zero"
        # metrics, J, medium: FP64 on both sides, rounded once to FP32
        err = np.abs(a - b)
        # one FP32 rounding of each value, plus a floor at 1e-12 of the component's scale
        cmax = np.abs(b).reshape(13, -1).max(1)
        cmax[:9] = cmax[:9].max()  # the nine metric components share one scale
        scale = cmax[:, None, None, None]
        tol = 2.5e-7 * np.abs(b) + 1e-12 * scale
        bad = int((err > tol).sum())
        assert bad == 0, f"terrain={use_terrain}: {bad} values off, worst excess {float((err - tol).max()):.3e}"
        g.close()


@pytest.mark.parametrize("fs,pml", [(1, 1), (0, 1), (1, 0)])
def test_steps_random_state_fp32(gpu, oracle_lib, fs, pml):
    cg, po = gpu, oracle_lib
    NX, NY, NZ = 34, 30, 26
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=8, nPML=6, free_surface=fs, pml=pml)
    g, o64 = _pair(cg, po, d, po.F64)
    o32 = po.Oracle(d, po.F32)
    o32.set_model(None, *models.homogeneous(NX, NY, NZ))
    w0 = models.smooth_state(NX, NY, NZ, seed=3)
    g.set_state(w0.astype(np.float32))
    o64.set_state(w0.astype(np.float32).astype(np.float64))
    o32.set_state(w0.astype(np.float32).astype(np.float64))
    for n in range(3):
        g.step(1)
        o64.step(1)
        o32.step(1)
        wg = g.state()
        e64 = rel_l2(wg, o64.state())
        e32 = rel_l2(wg, o32.state())
        assert e64 <= 1e-5, (n, e64)
        assert e32 <= 1e-5, (n, e32)


def test_point_source_stations_pgv_fp32(gpu, oracle_lib):
    cg, po = gpu, oracle_lib
    NX, NY, NZ = 48, 44, 36
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=120, nPML=8)
    g, o = _pair(cg, po, d, po.F64)
    st = models.surface_stations(NX, NY, NZ, n=4)
    for x in (g, o):
        x.set_source_point(NX // 2, NY // 2, NZ - 12, 2.0)
        x.set_stations(st)
    g.step(120)
    o.step(120)
    sg, so = g.stations(), o.stations()
    assert np.abs(so[:, :, :3]).max() > 0
    assert rel_l2(sg[:, :, :3], so[:, :, :3]) <= 1e-5
    assert rel_l2(sg, so) <= 1e-5
    assert rel_l2(g.pgv(), o.pgv()) <= 1e-5
    for axis, idx in ((0, NX // 2), (1, NY // 2), (2, NZ - 1)):
        sl = g.slice(axis, idx, 2)
        ref = o.state()[2]
        ref = {0: ref[:, :, idx], 1: ref[:, idx, :], 2: ref[idx]}[axis]
        assert rel_l2(sl, ref) <= 1e-5


def test_terrain_table_source_fp32(gpu, oracle_lib):
    cg, po = gpu, oracle_lib
    NX, NY, NZ = 40, 40, 30
    DH = 100.0
    d = cg.make_desc(NX, NY, NZ, DH=DH, DT=0.006, NT=100, nPML=8, use_terrain=1, depth=3000.0)
    terrain = models.gauss_hill(NX, NY, DH, h0=400.0, width=1200.0)
    z = models.terrain_z(terrain, NZ, 3000.0)
    med = models.layered_on_depth(z - terrain[3:-3, 3:-3][None], [0, 1.0, 2.0], [2000.0, 3000.0, 3500.0])
    g, o = _pair(cg, po, d, po.F64, terrain, med)
    nt, dt = 200, 0.004
    t = np.arange(nt) * dt
    fc = 2.5
    r = np.pi * fc * (t - 1.2 / fc)
    stf = (1 - 2 * r * r) * np.exp(-r * r)
    mt = po.moment_tensor(30.0, 70.0, 20.0)
    xyz = np.array([[20, 20, 15], [21, 20, 15], [20, 20, 15]], np.int32)  # one duplicate
    rate = np.zeros((nt, 3, 6), np.float32)
    for p in range(3):
        rate[:, p, :] = (1e15 / 3.0e10) * stf[:, None] * mt[None, :]
    st = models.surface_stations(NX, NY, NZ, n=3)
    for x in (g, o):
        x.set_source_table(xyz, rate, dt)
        x.set_stations(st)
    g.step(100)
    o.step(100)
    sg, so = g.stations(), o.stations()
    assert np.abs(so).max() > 0
    assert rel_l2(sg, so) <= 1e-5


@pytest.mark.parametrize("prec,tol", [("fp16", 1e-2), ("bf16", 5e-2)])
def test_low_precision_vs_fp64(gpu, oracle_lib, prec, tol):
    cg, po = gpu, oracle_lib
    NX, NY, NZ = 48, 44, 36
    d16 = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=150, nPML=8, precision=prec)
    d64 = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=150, nPML=8)
    med = models.homogeneous(NX, NY, NZ)
    g = cg.Domain(d16)
    g.set_model(None, *med)
    o = po.Oracle(d64, po.F64)
    o.set_model(None, *med)
    st = models.surface_stations(NX, NY, NZ, n=4)
    for x in (g, o):
        x.set_source_point(NX // 2, NY // 2, NZ - 12, 2.0)
        x.set_stations(st)
    g.step(150)
    o.step(150)
    e = rel_l2(g.stations()[:, :, :3], o.stations()[:, :, :3])
    print(f"{prec} seismogram rel L2 vs FP64 oracle: {e:.3e}")
    assert e <= tol


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_uniform_metric_pair_matches_per_point_reads(gpu, monkeypatch, prec):
    """On a terrain-following grid every interior point stores the same (xi_x, eta_y), and the
    stage kernels then take that pair from one stored copy: the wavefield after some steps is
    bit-for-bit the one of per-point reads (CGFDM_STREAM_PAIR0=1)."""
    cg = gpu
    NX, NY, NZ = 80, 44, 30
    d = cg.make_desc(NX, NY, NZ, DH=100.0, DT=0.006, NT=10, nPML=6, precision=prec, use_terrain=1, depth=3000.0)
    terrain = models.gauss_hill(NX, NY, 100.0, h0=400.0, width=1200.0)
    # storage units: 16-bit domains store velocity x 5e3 and stress x 1e-6
    scale = {"Cv": 5e3, "Cs": 1e-6} if prec == "fp16" else {}
    w0 = models.smooth_state(NX, NY, NZ, seed=5, **scale).astype(np.float32)
    states, flags = [], []
    for stream in ("0", "1"):
        monkeypatch.setenv("CGFDM_STREAM_PAIR0", stream)
        g = cg.Domain(d)
        g.set_model(terrain, *models.homogeneous(NX, NY, NZ))
        g.set_state(w0)
        g.step(4)
        states.append(g.state())
        flags.append(g.uniform_metric_pair())
        g.close()
    assert flags == [True, False]
    assert np.isfinite(states[0]).all() and np.abs(states[0]).max() > 0
    assert np.array_equal(states[0], states[1])
