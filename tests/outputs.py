"""Readers for the reference's per-rank output files and a restatement of its layered-medium
rule, used to check the command-line driver's files.

The readers are the library's own (cgfdm_outputs: partition init_grid.cpp:38-75,
pyscripts/GRID.py:10-93; layouts station.cpp:382-405, PGV.cpp:103-111, data_io.cpp:323-526), so
the driver tests exercise them on the files the driver writes.
"""
import struct

import numpy as np

from cgfdm_outputs import partition, read_pgv, read_plane, read_stations  # noqa: F401


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
