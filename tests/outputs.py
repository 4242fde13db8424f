"""Readers for the reference's per-rank output files and a restatement of its layered-medium
rule, used to check the command-line driver's files.

Partition: the first NX % PX ranks get one more point (init_grid.cpp:38-75,
pyscripts/GRID.py:10-93). File layouts: station.cpp:382-405, PGV.cpp:103-111, data_io.cpp:323-431.
"""
import os
import struct

import numpy as np


def partition(N, P):
    n = np.full(P, N // P)
    res = N % P
    n[:res] += 1
    front = np.concatenate([[0], np.cumsum(n)[:-1]])
    return n, front


def read_stations(out_dir, params, NT):
    """All station files -> (xyz[n,3], wave[n,NT,9]) in file order (rank order, then station order)."""
    xyz, wave = [], []
    for X in range(params["PX"]):
        for Y in range(params["PY"]):
            for Z in range(params["PZ"]):
                fn = os.path.join(out_dir, f"station_mpi_{X}_{Y}_{Z}.bin")
                if not os.path.exists(fn):
                    continue
                raw = open(fn, "rb").read()
                n = len(raw) // (3 * 4 + NT * 9 * 4)
                assert n * (12 + NT * 36) == len(raw), fn
                xyz.append(np.frombuffer(raw[: 12 * n], np.int32).reshape(n, 3))
                wave.append(np.frombuffer(raw[12 * n:], np.float32).reshape(n, NT, 9))
    return np.concatenate(xyz), np.concatenate(wave)


def read_pgv(out_dir, params):
    NX, NY = params["NX"], params["NY"]
    nx, fx = partition(NX, params["PX"])
    ny, fy = partition(NY, params["PY"])
    Z = params["PZ"] - 1
    out = np.zeros((NY, NX, 2), np.float32)
    for X in range(params["PX"]):
        for Y in range(params["PY"]):
            a = np.fromfile(os.path.join(out_dir, f"PGV_Z_mpi_{X}_{Y}_{Z}.bin"), np.float32)
            out[fy[Y]:fy[Y] + ny[Y], fx[X]:fx[X] + nx[X]] = a.reshape(ny[Y], nx[X], 2)
    return out


def read_plane(out_dir, name, axis, params):
    """Reassembles {name}_{X|Y|Z}_mpi_*.bin: axis 0 -> (NZ, NY), 1 -> (NZ, NX), 2 -> (NY, NX)."""
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
                fn = os.path.join(out_dir, f"{name}_{tag}_mpi_{X}_{Y}_{Z}.bin")
                if not os.path.exists(fn):
                    continue
                r = (X, Y, Z)
                nf, ff = parts[fast][0][r[fast]], parts[fast][1][r[fast]]
                ns, fs = parts[slow][0][r[slow]], parts[slow][1][r[slow]]
                out[fs:fs + ns, ff:ff + nf] = np.fromfile(fn, np.float32).reshape(ns, nf)
                found += 1
    assert found > 0, name
    return out


def write_layer_file(path, dep_m, vs, vp, rho):
    """The binary written by setLayeredModel.py: int32 layers, then per layer float32 depth, Vs,
    Vp, rho in km, km/s, km/s, g/cm^3."""
    with open(path, "wb") as f:
        f.write(struct.pack("i", len(dep_m)))
        for d, s, p, r in zip(dep_m, vs, vp, rho):
            f.write(struct.pack("ffff", d / 1000, s / 1000, p / 1000, r / 1000))


def layered_medium(coordz, terrain_surface, dep_m, vs, vp, rho, base=(3464.0, 6000.0, 2670.0), from_km=True):
    """Restatement of generate_1D_medium (medium.cpp:84-216) with LayeredStructureTerrain:
    the last matching rule among the layer tests wins; values start from the homogeneous default.
    All comparisons in float32, as the reference's floats. from_km: the values pass through the
    layer file's km units (float km * 1000, as the reference reads them); otherwise they are
    taken as float32 metres and m/s."""
    f32 = np.float32
    conv = (lambda v: f32(f32(v / 1000) * f32(1000.0))) if from_km else f32
    H = [conv(d) for d in dep_m]
    VS = [conv(v) for v in vs]
    VP = [conv(v) for v in vp]
    RH = [conv(v) for v in rho]
    cz = coordz.astype(np.float32)
    ter = np.broadcast_to(terrain_surface, cz.shape).astype(np.float64)
    out = [np.full(cz.shape, b, np.float32) for b in base]
    tol = f32(1.0)
    for F in range(1, len(H)):
        h0 = (-np.float64(H[0]) + ter).astype(np.float32)
        h1 = (-np.float64(H[F - 1]) + ter).astype(np.float32)
        h2 = (-np.float64(H[F]) + ter).astype(np.float32)
        for m, (a, b, c) in enumerate(((VS[0], VP[0], RH[0]), (VS[F - 1], VP[F - 1], RH[F - 1]), (VS[F], VP[F], RH[F]))):
            if m == 0:
                sel = cz > h0 - tol
            elif m == 1:
                sel = (h1 - tol > cz) & (cz > h2 - tol)
            else:
                sel = h2 - tol > cz
            out[0] = np.where(sel, a, out[0])
            out[1] = np.where(sel, b, out[1])
            out[2] = np.where(sel, c, out[2])
    return out
