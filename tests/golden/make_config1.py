"""Writes tests/golden/config1_fp64.npz: the FP64 oracle's output for the half-space workload at
its full size, the fixture tests/test_gpu_configs.py compares the HIP path with.

Workload: flat free surface, homogeneous Vp 6000 m/s, Vs 3464 m/s, rho 2670 kg/m^3 (the
reference's default medium, medium.cpp:75-77), 201 x 201 x 101 points, DH 100 m, DT 0.008 s,
1000 steps, 12-cell PML on the five non-surface faces, explosive Ricker point source
(M0 = 1e16 N m, fc 2 Hz) at the centre 2 km below the surface (X = Y = 100, Z = 80), and a
21 x 21 surface station grid. Stored: Vx, Vy, Vz of every other station of that grid in x and y
(11 x 11 = 121 stations, all 1000 steps) and the whole PGV map, float32.

The oracle is the checker (oracle/oracle.cpp, FP64); this script runs it on the CPU, which
takes about half an hour on 8 cores. Usage: python tests/golden/make_config1.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "cgfdm-fp16_sve_sme_amd")):
    sys.path.insert(0, p)

import cgfdm  # noqa: E402  (descriptor structure only; no device call)
import models  # noqa: E402
import pyoracle  # noqa: E402
import workloads  # noqa: E402

NX, NY, NZ = 201, 201, 101
NT = 1000
SRC = (100, 100, 80)
FC = 2.0


def stations():
    """The 21 x 21 surface grid and the indices of the stored subset (every other in x and y)."""
    st = workloads.surface_stations(NX, NY, NZ, n=21)
    sel = np.array([r * 21 + c for r in range(0, 21, 2) for c in range(0, 21, 2)])
    return st, sel


def desc():
    return cgfdm.make_desc(NX, NY, NZ, DH=100.0, DT=0.008, NT=NT, nPML=12)


def main():
    st, sel = stations()
    o = pyoracle.Oracle(desc(), pyoracle.F64)
    o.set_model(None, *models.homogeneous(NX, NY, NZ))
    o.set_source_point(*SRC, FC)
    o.set_stations(st)
    t0 = time.time()
    for n in range(0, NT, 50):
        o.step(50)
        print(f"step {n + 50}/{NT}  {time.time() - t0:.0f} s", flush=True)
    s = o.stations()
    out = os.path.join(HERE, "config1_fp64.npz")
    np.savez_compressed(out, stations=s[sel, :, :3].astype(np.float32), pgv=o.pgv().astype(np.float32),
                        station_xyz=st[sel], nt=NT)
    print("wrote", out)


if __name__ == "__main__":
    main()
