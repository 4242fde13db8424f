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
