"""Output tooling (cgfdm_outputs): rank partition, readers of the reference's per-rank files and
the comparison scripts' error measures, checked on the CPU against closed forms and against files
laid out as the reference's writers lay them out (station.cpp:382-405, PGV.cpp:103-111,
data_io.cpp:323-526). The driver tests (test_cli.py, test_gpu_cli.py) read the driver's own files
with the same readers."""
import json

import numpy as np
import pytest

import cgfdm_outputs as co


def test_partition_matches_grid_rule():
    # GRID.py:34-60: the first N % P ranks get one more point
    n, f = co.partition(10, 3)
    assert n.tolist() == [4, 3, 3] and f.tolist() == [0, 4, 7]
    n, f = co.partition(401, 4)
    assert n.tolist() == [101, 100, 100, 100] and f.tolist() == [0, 101, 201, 301]
    n, f = co.partition(7, 7)
    assert n.tolist() == [1] * 7 and f.tolist() == list(range(7))
    with pytest.raises(ValueError):
        co.partition(3, 4)


def test_station_nt_uses_float32_dt():
    # 1000 steps for TMAX = 8.0001 at DT = 0.008; 8 / float32(0.008) = 999.99995 -> 999
    assert co.station_nt({"TMAX": 8.0001, "DT": 0.008}) == 1000
    assert co.station_nt({"TMAX": 8.0, "DT": 0.008}) == 999
    assert co.station_nt({"TMAX": 1.0, "DT": 0.5}) == 2


def _params(NX=11, NY=9, NZ=7, PX=2, PY=3, PZ=2):
    return {"NX": NX, "NY": NY, "NZ": NZ, "PX": PX, "PY": PY, "PZ": PZ, "TMAX": 0.05, "DT": 0.01}


def _write_run(d, params, field, pgv, st_xyz, st_wave):
    """Per-rank files of a run: the Z plane of `field` at the top ranks, X/Y planes at given
    indices, PGV on the top ranks, stations on the rank that owns them."""
    d.mkdir(parents=True, exist_ok=True)
    P = params
    parts = [co.partition(P["NX"], P["PX"]), co.partition(P["NY"], P["PY"]), co.partition(P["NZ"], P["PZ"])]
    for X in range(P["PX"]):
        for Y in range(P["PY"]):
            for Z in range(P["PZ"]):
                (nx, fx), (ny, fy), (nz, fz) = [(parts[a][0][r], parts[a][1][r]) for a, r in enumerate((X, Y, Z))]
                tag = f"_mpi_{X}_{Y}_{Z}.bin"
                if Z == P["PZ"] - 1:
                    field[-1, fy:fy + ny, fx:fx + nx].astype(np.float32).tofile(str(d / ("FreeSurfVz_3_Z" + tag)))
                    pgv[fy:fy + ny, fx:fx + nx].astype(np.float32).tofile(str(d / ("PGV_Z" + tag)))
                if fx <= 4 < fx + nx:  # X plane at global x = 4: [nz][ny]
                    field[fz:fz + nz, fy:fy + ny, 4].astype(np.float32).tofile(str(d / ("Vz_3_X" + tag)))
                if fy <= 2 < fy + ny:  # Y plane at global y = 2: [nz][nx]
                    field[fz:fz + nz, 2, fx:fx + nx].astype(np.float32).tofile(str(d / ("Vz_3_Y" + tag)))
                own = [s for s in range(len(st_xyz)) if fx <= st_xyz[s, 0] < fx + nx and fy <= st_xyz[s, 1] < fy + ny
                       and fz <= st_xyz[s, 2] < fz + nz]
                if own:
                    with open(d / ("station" + tag), "wb") as f:
                        st_xyz[own].astype(np.int32).tofile(f)
                        st_wave[own].astype(np.float32).tofile(f)


def _run(rng, params):
    P = params
    field = rng.standard_normal((P["NZ"], P["NY"], P["NX"])).astype(np.float32)
    pgv = np.abs(rng.standard_normal((P["NY"], P["NX"], 2))).astype(np.float32)
    xyz = np.array([[1, 1, 6], [9, 7, 6], [5, 4, 2]], np.int32)
    wave = rng.standard_normal((3, co.station_nt(P), 9)).astype(np.float32)
    return field, pgv, xyz, wave


def test_readers_reassemble_rank_files(tmp_path):
    P = _params()
    field, pgv, xyz, wave = _run(np.random.default_rng(1), P)
    _write_run(tmp_path / "a", P, field, pgv, xyz, wave)
    np.testing.assert_array_equal(co.read_plane(tmp_path / "a", "FreeSurfVz_3", 2, P), field[-1])
    np.testing.assert_array_equal(co.read_plane(tmp_path / "a", "Vz_3", 0, P), field[:, :, 4])
    np.testing.assert_array_equal(co.read_plane(tmp_path / "a", "Vz_3", 1, P), field[:, 2, :])
    np.testing.assert_array_equal(co.read_pgv(tmp_path / "a", P), pgv)
    rx, rw = co.read_stations(tmp_path / "a", P)
    order = [np.flatnonzero((xyz == r).all(1))[0] for r in rx]
    assert sorted(order) == [0, 1, 2]
    np.testing.assert_array_equal(rw, wave[order])
    with pytest.raises(FileNotFoundError):
        co.read_plane(tmp_path / "a", "Vx_3", 2, P)
    with pytest.raises(ValueError):
        co.read_stations(tmp_path / "a", P, NT=co.station_nt(P) + 1)


def test_relative_error_percent_closed_form():
    err, mx = co.relative_error_percent([1.0, -3.0, 2.0], [1.0, -4.0, 2.0])
    np.testing.assert_allclose(err, [0.0, 25.0, 0.0])
    assert mx == pytest.approx(25.0)
    with pytest.raises(ValueError):
        co.relative_error_percent([1.0], [0.0])


def test_seismogram_errors():
    std = np.zeros((2, 4, 9))
    std[1, 2, 3] = -5.0
    test = std.copy()
    test[0, 1, 0] = 0.5
    diff, vmax = co.seismogram_errors(test, std)
    assert vmax == 5.0 and diff[0, 1, 0] == 0.5 and np.count_nonzero(diff) == 1
    with pytest.raises(ValueError):
        co.seismogram_errors(test[:1], std)


def test_intensity_table_edges():
    # plotPGV.py:123-146: lower bounds inclusive
    edges = [0.001, 0.005, 0.01, 0.02, 0.05, 0.10, 0.19, 0.36, 0.72, 1.41]
    vals = [0.0, 0.0009] + edges + [5.0]
    want = [1, 1] + list(range(2, 12)) + [11]
    assert co.intensity(vals).tolist() == want
    below = [np.nextafter(e, 0.0) for e in edges]
    assert co.intensity(below).tolist() == list(range(1, 11))


def test_compare_two_runs(tmp_path, capsys):
    P = _params()
    rng = np.random.default_rng(2)
    field, pgv, xyz, wave = _run(rng, P)
    _write_run(tmp_path / "std", P, field, pgv, xyz, wave)
    f2, w2 = field.copy(), wave.copy()
    vm = np.abs(field[-1]).max()
    f2[-1, 3, 5] += 0.02 * vm  # 2 % of the largest surface value
    w2[1, 2, 0] += 0.5
    p2 = pgv.copy()
    p2[0, 0, 0] = 2.0  # intensity 11 where the standard has less
    pgv[0, 0, 0] = 0.3
    _write_run(tmp_path / "std", P, field, pgv, xyz, wave)
    _write_run(tmp_path / "test", P, f2, p2, xyz, w2)
    out = co.compare(tmp_path / "std", tmp_path / "test", P, var="Vz", it=3)
    assert out["plane"]["max_relative_error_percent"] == pytest.approx(2.0, rel=1e-5)
    assert out["stations"]["n"] == 3 and out["stations"]["max_abs_difference"] == pytest.approx(0.5, rel=1e-6)
    assert out["pgv"]["intensity_cells_differing"] == 1
    pfile = tmp_path / "params.json"
    pfile.write_text(json.dumps(P))
    assert co.main(["compare", str(tmp_path / "std"), str(tmp_path / "test"), "--params", str(pfile),
                    "--it", "3"]) == 0
    printed = json.loads(capsys.readouterr().out)
    assert printed["plane"]["name"] == "FreeSurfVz_3"
