"""The command-line driver end to end on the GPU: a Gauss-hill run split over a 2 x 2 x 1 rank
grid; the per-rank station, PGV and snapshot files, reassembled as the reference's scripts do
(pyscripts/GRID.py), must match the FP64 oracle on the same model to 1e-5 relative L2."""
import json
import os
import subprocess

import numpy as np
import pytest

import models
import outputs
from models import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cgfdm-fp16_sve_sme_amd", "cgfdm_run")


def test_cli_matches_oracle(tmp_path, gpu, oracle_lib):
    po = oracle_lib
    NX, NY, NZ, DH, DT = 40, 36, 30, 100.0, 0.008
    nsteps = 50
    params = {"TMAX": nsteps * DT + 1e-4, "DT": DT, "DH": DH, "NX": NX, "NY": NY, "NZ": NZ,
              "PX": 2, "PY": 2, "PZ": 1, "centerX": NX // 2, "centerY": NY // 2, "IT_SKIP": 25,
              "sliceX": 10, "sliceY": 12, "sliceZ": 20, "sliceFreeSurf": 1, "nPML": 8,
              "gauss_hill": 1, "useTerrain": 0, "useMedium": 0, "useMultiSource": 0,
              "useSingleSource(ricker)": 1, "rickerfc": 2.0, "sourceX": 20, "sourceY": 18,
              "sourceZ": 20, "out": "out", "Depth(km)": 9.9}
    (tmp_path / "params.json").write_text(json.dumps(params))
    st = [[5, 5, NZ - 1], [20, 18, NZ - 1], [33, 30, 25], [12, 25, NZ - 1]]
    (tmp_path / "station.json").write_text(json.dumps({"station(point)": {f"S{q}": s for q, s in enumerate(st)}}))
    r = subprocess.run([CLI], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = str(tmp_path / "out")
    NT = int(params["TMAX"] / np.float32(DT))
    assert NT == nsteps

    # the same model through the oracle
    D = DH * NZ
    x = (np.arange(NX + 6) - 3 - NX // 2) * DH
    y = (np.arange(NY + 6) - 3 - NY // 2) * DH
    X, Y = np.meshgrid(x, y)
    terrain = 0.2 * D * np.exp(-0.5 * (X ** 2 + Y ** 2) / (0.1 * D) ** 2)
    import cgfdm
    d = cgfdm.make_desc(NX, NY, NZ, DH=DH, DT=DT, NT=NT, nPML=8, use_terrain=1, depth=D)
    o = po.Oracle(d, po.F64)
    o.set_model(terrain, *models.homogeneous(NX, NY, NZ))
    o.set_source_point(20, 18, 20, 2.0)
    o.set_stations(np.array(st, np.int32))
    snaps = {}
    for it in range(NT):
        o.step(1)
        if it % params["IT_SKIP"] == 0:
            snaps[it] = o.state()[:3]
    xyz, wave = outputs.read_stations(out, params, NT)
    # file order is rank order; map back to the station list
    order = [st.index(list(p)) for p in xyz]
    ref = o.stations()[order]
    assert rel_l2(wave, ref) <= 1e-5
    assert rel_l2(outputs.read_pgv(out, params), o.pgv()) <= 1e-5
    for it, w in snaps.items():
        vz = outputs.read_plane(out, f"Vz_{it}", 2, params)
        assert rel_l2(vz, w[2, params["sliceZ"]]) <= 1e-5
        vx = outputs.read_plane(out, f"Vx_{it}", 0, params)
        assert rel_l2(vx, w[0][:, :, params["sliceX"]]) <= 1e-5
        fs = outputs.read_plane(out, f"FreeSurfVy_{it}", 2, params)
        assert rel_l2(fs, w[1, NZ - 1]) <= 1e-5


def test_cli_finite_fault_matches_oracle(tmp_path, gpu, oracle_lib):
    """useMultiSource: a finite-fault file (reference layout) on a Gauss hill, mapped by the
    driver (cgfdm_finite_fault_map) and by the oracle's restatement of multiSource.cpp; the
    stations must match the FP64 oracle driven by the oracle's table to 1e-5 relative L2."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import finite_fault as ff
    po = oracle_lib
    NX, NY, NZ, DH, DT = 40, 36, 30, 100.0, 0.008
    nsteps = 60
    lon0, lat0 = 101.2, 29.6
    params = {"TMAX": nsteps * DT + 1e-4, "DT": DT, "DH": DH, "NX": NX, "NY": NY, "NZ": NZ,
              "PX": 1, "PY": 2, "PZ": 1, "centerX": NX // 2, "centerY": NY // 2, "IT_SKIP": 1000,
              "sliceX": 10, "sliceY": 12, "sliceZ": 20, "sliceFreeSurf": 0, "nPML": 8,
              "gauss_hill": 1, "useTerrain": 0, "useMedium": 0, "useMultiSource": 1,
              "useSingleSource(ricker)": 0, "sourceFile": "fault.bin", "sourceDir": ".",
              "centerLongitude": lon0, "centerLatitude": lat0, "out": "out", "Depth(km)": 9.9}
    (tmp_path / "params.json").write_text(json.dumps(params))
    st = [[5, 5, NZ - 1], [20, 18, NZ - 1], [33, 30, 25], [12, 25, NZ - 1]]
    (tmp_path / "station.json").write_text(json.dumps({"station(point)": {f"S{q}": s for q, s in enumerate(st)}}))
    # a 6 x 4 patch of subfaults (strike 30, dip 60) around 1.2 km depth, 1 Hz triangles
    rng = np.random.default_rng(21)
    npts, nt, sdt = 24, 40, 0.01
    R = 6371e3
    u, v = np.meshgrid(np.arange(6) * 150.0 - 375.0, np.arange(4) * 120.0 - 180.0)
    lon = lon0 + np.degrees((u.ravel() + 40.0) / (R * np.cos(np.radians(lat0))))
    lat = lat0 + np.degrees((0.6 * u.ravel() - 30.0) / R)
    z = -1200.0 + v.ravel()
    t = np.arange(nt) * sdt
    rate = np.array([np.clip(1 - np.abs(t - 0.1 - 0.01 * q) / 0.1, 0, None) for q in range(npts)]) * 1e-1
    rake = np.tile(rng.uniform(60, 120, (npts, 1)), (1, nt))
    path = tmp_path / "fault.bin"
    ff.write_fault(path, lon, lat, z, np.full(npts, 1.8e4), np.full(npts, 30.0), np.full(npts, 60.0), rake, rate, sdt)
    r = subprocess.run([CLI], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = str(tmp_path / "out")
    NT = int(params["TMAX"] / np.float32(DT))

    D = DH * NZ
    x = (np.arange(NX + 6) - 3 - NX // 2) * DH
    y = (np.arange(NY + 6) - 3 - NY // 2) * DH
    X, Y = np.meshgrid(x, y)
    terrain = 0.2 * D * np.exp(-0.5 * (X ** 2 + Y ** 2) / (0.1 * D) ** 2)
    import cgfdm
    d = cgfdm.make_desc(NX, NY, NZ, DH=DH, DT=DT, NT=NT, nPML=8, use_terrain=1, depth=D)
    xyz, table, tdt = ff.finite_fault_map(path, d, terrain, lon0, lat0)
    assert len(xyz) > 4
    o = po.Oracle(d, po.F64)
    o.set_model(terrain, *models.homogeneous(NX, NY, NZ))
    o.set_source_table(xyz, table, tdt)
    o.set_stations(np.array(st, np.int32))
    o.step(NT)
    sxyz, wave = outputs.read_stations(out, params, NT)
    order = [st.index(list(p)) for p in sxyz]
    ref = o.stations()[order]
    assert np.abs(ref).max() > 0
    assert rel_l2(wave, ref) <= 1e-5


def test_cli_output_compare_fp16_vs_fp32(tmp_path, gpu):
    """The output tooling on the driver's own files: the same Gauss-hill run with FP32 and FP16
    storage, compared as the reference's comparison scripts do (cgfdm_outputs.compare): zero
    against itself, and FP16 within a few percent of FP32 pointwise (of the largest value)."""
    import cgfdm_outputs as co
    NX, NY, NZ, DH, DT = 40, 36, 30, 100.0, 0.008
    nsteps = 60
    params = {"TMAX": nsteps * DT + 1e-4, "DT": DT, "DH": DH, "NX": NX, "NY": NY, "NZ": NZ,
              "PX": 2, "PY": 1, "PZ": 1, "centerX": NX // 2, "centerY": NY // 2, "IT_SKIP": 50,
              "sliceX": 10, "sliceY": 12, "sliceZ": 20, "sliceFreeSurf": 1, "nPML": 8,
              "gauss_hill": 1, "useTerrain": 0, "useMedium": 0, "useMultiSource": 0,
              "useSingleSource(ricker)": 1, "rickerfc": 2.0, "sourceX": 20, "sourceY": 18,
              "sourceZ": 24, "Depth(km)": 9.9}
    st = [[5, 5, NZ - 1], [20, 18, NZ - 1], [33, 30, NZ - 1], [12, 25, NZ - 1]]
    (tmp_path / "station.json").write_text(json.dumps({"station(point)": {f"S{q}": s for q, s in enumerate(st)}}))
    for prec in ("fp32", "fp16"):
        (tmp_path / "params.json").write_text(json.dumps(dict(params, out=f"out_{prec}")))
        r = subprocess.run([CLI, "--precision", prec], cwd=tmp_path, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
    a, b = str(tmp_path / "out_fp32"), str(tmp_path / "out_fp16")
    same = co.compare(a, a, params, var="Vz", it=50)
    assert same["plane"]["max_relative_error_percent"] == 0.0
    assert same["stations"]["max_abs_difference"] == 0.0 and same["pgv"]["intensity_cells_differing"] == 0
    res = co.compare(a, b, params, var="Vz", it=50)
    assert res["plane"]["max_abs_standard"] > 0 and res["stations"]["n"] == len(st)
    assert 0 < res["plane"]["max_relative_error_percent"] < 2.0
    assert 0 < res["stations"]["max_relative_error_percent"] < 2.0
    assert res["pgv"]["max_relative_error_percent"] < 2.0
