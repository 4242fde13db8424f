"""Command-line driver (cgfdm_run): params.json / station.json / layer-file handling and the model
planes it writes, checked without a GPU (--model-only stops before any device call)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import outputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "cgfdm-fp16_sve_sme_amd", "cgfdm_run")
GOLD = os.path.join(ROOT, "tests", "golden")


def _halfspace_layers():
    """HalfSpace.dat: depth, rho, Vp, Vs columns (read by setLayeredModel.py:41-58 with Trans = 1)."""
    d = np.loadtxt(os.path.join(GOLD, "HalfSpace.dat")).T
    return d[0], d[3], d[2], d[1]


def _case(tmp_path, **over):
    params = json.load(open(os.path.join(GOLD, "halfspace_params.json")))
    params.update({"NX": 30, "NY": 26, "NZ": 250, "PX": 2, "PY": 2, "PZ": 3, "sliceX": 11, "sliceY": 7,
                   "sliceZ": 240, "nPML": 8, "out": "out", "LayeredFileName": "layers.bin"})
    params.update(over)
    (tmp_path / "params.json").write_text(json.dumps(params))
    shutil.copy(os.path.join(GOLD, "station.json"), tmp_path / "station.json")
    dep, vs, vp, rho = _halfspace_layers()
    outputs.write_layer_file(str(tmp_path / "layers.bin"), dep, vs, vp, rho)
    return params


def _run(tmp_path, *args):
    r = subprocess.run([CLI, *args], cwd=tmp_path, capture_output=True, text=True)
    return r


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CLI):
        pytest.fail("cgfdm_run is not built (run __graft_entry__.build())")


def test_layered_medium_flat_known_answer(tmp_path):
    params = _case(tmp_path)
    r = _run(tmp_path, "--model-only")
    assert r.returncode == 0, r.stderr
    out = str(tmp_path / "out")
    NZ, DH = params["NZ"], params["DH"]
    vs = outputs.read_plane(out, "Vs", 0, params)  # (NZ, NY) at X = sliceX
    z = (np.arange(NZ) + 1) * DH - NZ * DH
    # the rule of medium.cpp:163-206 applied to HalfSpace.dat: 0..10 km -> 2500, then 3000 to
    # 15 km, 3500 to 20 km, 4000 below (the first row's depth is the top of the second layer's
    # test, so the first layer reaches the second depth)
    expect = np.select([z > -10001, z > -15001, z > -20001], [2500.0, 3000.0, 3500.0], 4000.0)
    np.testing.assert_array_equal(vs[:, 5], expect.astype(np.float32))
    zc = outputs.read_plane(out, "coordZ", 1, params)
    np.testing.assert_allclose(zc[:, 3], z, rtol=0, atol=1e-3)


def test_layered_medium_follows_gauss_hill(tmp_path):
    params = _case(tmp_path, gauss_hill=1, NZ=60, sliceZ=50, PZ=1)
    r = _run(tmp_path, "--model-only")
    assert r.returncode == 0, r.stderr
    out = str(tmp_path / "out")
    NX, NY, NZ, DH = params["NX"], params["NY"], params["NZ"], params["DH"]
    D = DH * NZ  # Depth is reset to DH*NZ without useTerrain (getParams.cpp:244-247)
    x = (np.arange(NX + 6) - 3 - params["centerX"]) * DH
    y = (np.arange(NY + 6) - 3 - params["centerY"]) * DH
    X, Y = np.meshgrid(x, y)
    ter = 0.2 * D * np.exp(-0.5 * (X ** 2 + Y ** 2) / (0.1 * D) ** 2)  # coord.cpp:164-180
    sy = params["sliceY"]
    K = np.arange(NZ)[:, None]
    h = ter[sy + 3, 3:-3][None, :]
    cz = -D + (h + D) / (NZ - 1) * K
    zc = outputs.read_plane(out, "coordZ", 1, params)
    np.testing.assert_allclose(zc, cz, rtol=0, atol=2e-3 * D)
    dep, vs_, vp_, rho_ = _halfspace_layers()
    ref = outputs.layered_medium(cz, h, dep, vs_, vp_, rho_)
    for name, a in zip(("Vs", "Vp", "rho"), ref):
        np.testing.assert_array_equal(outputs.read_plane(out, name, 1, params), a)


def test_bad_inputs_exit_codes(tmp_path):
    (tmp_path / "params.json").write_text("{ not json")
    assert _run(tmp_path, "--model-only").returncode == 101
    _case(tmp_path, **{"useSingleSource(ricker)": 0})
    assert _run(tmp_path, "--model-only").returncode == 180
    _case(tmp_path, useTerrain=1)
    assert _run(tmp_path, "--model-only").returncode == 1
    os.remove(tmp_path / "params.json")
    r = _run(tmp_path, "--model-only")
    assert r.returncode == 100 and "There is not" in r.stderr
