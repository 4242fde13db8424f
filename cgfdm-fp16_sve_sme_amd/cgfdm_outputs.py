"""Readers for the per-rank output files of the command-line driver (`cgfdm_run`), which writes
them in the reference's layouts, and the error measures of the reference's comparison scripts,
so that runs of this library and of the reference can be compared file for file.

Rank partition (init_grid.cpp:38-75, pyscripts/GRID.py:10-93): along an axis of N points over P
ranks, the first N % P ranks hold N // P + 1 points and the others N // P, in rank order.

File layouts, one file per rank `..._mpi_X_Y_Z.bin`, native-endian:
  station_mpi_X_Y_Z.bin      int32 XYZ[n][3], then float32 wave[n][NT][9] (station.cpp:382-405),
                             NT = int(TMAX / float32(DT)) (propagate.cpp:35-36)
  PGV_Z_mpi_X_Y_Z.bin        float32 [ny][nx][2] (PGVh, PGV) on the top ranks (PGV.cpp:103-111)
  {name}_{X|Y|Z}_mpi_X_Y_Z.bin  one float32 plane on each rank that holds it (data_io.cpp:323-526):
                             X planes [nz][ny], Y planes [nz][nx], Z planes [ny][nx]; name is for
                             example Vz_800, FreeSurfVz_800, coordX, Vs

Error measures:
  relative_error_percent     |test - standard| / max|standard| x 100, pointwise (plotXY.py:128-145)
  seismogram_errors          station traces minus the standard's, and the largest standard
                             amplitude they are plotted against (plotMutiSeismos.py:139-159)
  intensity                  seismic intensity 1-11 from PGV in m/s (plotPGV.py:123-146)

Command line (prints the measures of two output folders):
    python cgfdm_outputs.py compare STANDARD_DIR TEST_DIR --params params.json [--var Vz --it 800]
"""
import argparse
import json
import os
import sys

import numpy as np

__all__ = ["partition", "station_nt", "read_stations", "read_pgv", "read_plane", "relative_error_percent",
           "seismogram_errors", "intensity", "compare"]


def partition(N, P):
    """(points per rank, first global index per rank) of N points over P ranks."""
    if P < 1 or N < P:
        raise ValueError(f"cannot split {N} points over {P} ranks")
    n = np.full(P, N // P, np.int64)
    res = N % P
    n[:res] += 1
    front = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
    return n, front


def station_nt(params):
    """Samples per station trace: int(TMAX / DT) with DT rounded to float32 first."""
    return int(float(params["TMAX"]) / float(np.float32(params["DT"])))


def _rank_file(out_dir, stem, X, Y, Z):
    return os.path.join(out_dir, f"{stem}_mpi_{X}_{Y}_{Z}.bin")


def read_stations(out_dir, params, NT=None):
    """All station files -> (xyz [n, 3] int32 global grid indices, wave [n, NT, 9] float32), in
    file order (ranks X, then Y, then Z; stations in each file's order)."""
    NT = station_nt(params) if NT is None else NT
    rec = 12 + NT * 36
    xyz, wave = [], []
    for X in range(params["PX"]):
        for Y in range(params["PY"]):
            for Z in range(params["PZ"]):
                fn = _rank_file(out_dir, "station", X, Y, Z)
                if not os.path.exists(fn):
                    continue
                raw = np.fromfile(fn, np.uint8)
                if raw.size % rec:
                    raise ValueError(f"{fn}: {raw.size} bytes is not a whole number of {NT}-sample stations")
                n = raw.size // rec
                xyz.append(raw[:12 * n].view(np.int32).reshape(n, 3))
                wave.append(raw[12 * n:].view(np.float32).reshape(n, NT, 9))
    if not xyz:
        raise FileNotFoundError(f"no station files in {out_dir}")
    return np.concatenate(xyz), np.concatenate(wave)


def read_pgv(out_dir, params):
    """PGV files of the top ranks -> [NY, NX, 2] float32 (PGVh, PGV)."""
    NX, NY = params["NX"], params["NY"]
    nx, fx = partition(NX, params["PX"])
    ny, fy = partition(NY, params["PY"])
    Z = params["PZ"] - 1
    out = np.zeros((NY, NX, 2), np.float32)
    for X in range(params["PX"]):
        for Y in range(params["PY"]):
            a = np.fromfile(_rank_file(out_dir, "PGV_Z", X, Y, Z), np.float32)
            out[fy[Y]:fy[Y] + ny[Y], fx[X]:fx[X] + nx[X]] = a.reshape(ny[Y], nx[X], 2)
    return out


def read_plane(out_dir, name, axis, params):
    """Reassembles {name}_{X|Y|Z}_mpi_*.bin over the ranks that wrote it: axis 0 -> [NZ, NY],
    1 -> [NZ, NX], 2 -> [NY, NX]; points no file covers are NaN."""
    NX, NY, NZ = params["NX"], params["NY"], params["NZ"]
    parts = [partition(NX, params["PX"]), partition(NY, params["PY"]), partition(NZ, params["PZ"])]
    fast = 1 if axis == 0 else 0
    slow = 1 if axis == 2 else 2
    shape = {0: (NZ, NY), 1: (NZ, NX), 2: (NY, NX)}[axis]
    out = np.full(shape, np.nan, np.float32)
    tag = "XYZ"[axis]
    found = 0
    for X in range(params["PX"]):
        for Y in range(params["PY"]):
            for Z in range(params["PZ"]):
                fn = _rank_file(out_dir, f"{name}_{tag}", X, Y, Z)
                if not os.path.exists(fn):
                    continue
                r = (X, Y, Z)
                nf, ff = parts[fast][0][r[fast]], parts[fast][1][r[fast]]
                ns, fs = parts[slow][0][r[slow]], parts[slow][1][r[slow]]
                out[fs:fs + ns, ff:ff + nf] = np.fromfile(fn, np.float32).reshape(ns, nf)
                found += 1
    if not found:
        raise FileNotFoundError(f"no {name}_{tag}_mpi_*.bin files in {out_dir}")
    return out


def relative_error_percent(test, standard):
    """Pointwise |test - standard| / max|standard| x 100, and its maximum."""
    test = np.asarray(test, np.float64)
    standard = np.asarray(standard, np.float64)
    vm = np.abs(standard).max()
    if not vm > 0:
        raise ValueError("the standard field is zero")
    err = np.abs(test - standard) / vm * 100.0
    return err, float(err.max())


def seismogram_errors(test, standard):
    """(test - standard, the largest |standard| over all stations and samples) for traces of
    the same stations ([n, NT, ...] arrays)."""
    test = np.asarray(test, np.float64)
    standard = np.asarray(standard, np.float64)
    if test.shape != standard.shape:
        raise ValueError(f"trace shapes differ: {test.shape} vs {standard.shape}")
    return test - standard, float(np.abs(standard).max())


# lower PGV bound (m/s) of intensities 2 .. 11
_INTENSITY_EDGES = np.array([0.001, 0.005, 0.01, 0.02, 0.05, 0.10, 0.19, 0.36, 0.72, 1.41])


def intensity(pgv):
    """Seismic intensity 1-11 of PGV values in m/s: 1 below 0.001, then one step at each of
    0.001, 0.005, 0.01, 0.02, 0.05, 0.10, 0.19, 0.36, 0.72 and 1.41 (lower bounds inclusive)."""
    return (1 + np.searchsorted(_INTENSITY_EDGES, np.asarray(pgv, np.float64), side="right")).astype(np.int32)


def compare(standard_dir, test_dir, params, var="Vz", it=None):
    """The comparison scripts' measures for two output folders of the same configuration:
    free-surface plane of `var` at step `it` (when written), station traces and PGV."""
    out = {}
    if it is not None:
        name = f"FreeSurf{var}_{it}"
        a = read_plane(standard_dir, name, 2, params)
        b = read_plane(test_dir, name, 2, params)
        out["plane"] = {"name": name, "max_relative_error_percent": relative_error_percent(b, a)[1],
                        "max_abs_standard": float(np.abs(a).max())}
    try:
        xa, wa = read_stations(standard_dir, params)
        xb, wb = read_stations(test_dir, params)
    except FileNotFoundError:
        xa = None
    if xa is not None:
        if not np.array_equal(xa, xb):
            raise ValueError("the two runs hold different stations")
        diff, vmax = seismogram_errors(wb, wa)
        out["stations"] = {"n": int(len(xa)), "max_abs_standard": vmax,
                           "max_abs_difference": float(np.abs(diff).max()),
                           "max_relative_error_percent": float(np.abs(diff).max() / vmax * 100.0) if vmax > 0 else None}
    try:
        pa, pb = read_pgv(standard_dir, params), read_pgv(test_dir, params)
    except FileNotFoundError:
        pa = None
    if pa is not None:
        out["pgv"] = {"max_relative_error_percent": relative_error_percent(pb[..., 0], pa[..., 0])[1],
                      "intensity_cells_differing": int((intensity(pb[..., 0]) != intensity(pa[..., 0])).sum())}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Compare two CGFDM output folders (reference layouts).")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("compare")
    c.add_argument("standard")
    c.add_argument("test")
    c.add_argument("--params", required=True)
    c.add_argument("--var", default="Vz")
    c.add_argument("--it", type=int, default=None)
    a = ap.parse_args(argv)
    with open(a.params) as f:
        params = json.load(f)
    print(json.dumps(compare(a.standard, a.test, params, a.var, a.it), indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
