"""Writes tests/golden/config5_reduced_fp64.npz: the FP64 oracle's output for the finite-fault
workload at 500 m spacing (workloads.finite_fault_workload(DH=500): 241 x 241 x 81 points,
DT 0.04 s, 800 steps = 32 s, fractal terrain, perturbed layered medium, 80 x 20 subfault rupture
written as a reference-layout source file and mapped by the oracle's restatement of
multiSource.cpp), the fixture tests/test_gpu_configs.py compares the HIP path with. Stored:
Vx, Vy, Vz of every other station of the 20 x 20 surface grid in x and y (100 stations, all
steps), the whole PGV map and the mapped source nodes, float32.

The oracle is the checker (oracle/oracle.cpp, FP64); this script runs it on the CPU, about a
quarter of an hour on 8 cores. Usage: python tests/golden/make_config5_reduced.py
"""
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests"), os.path.join(ROOT, "cgfdm-fp16_sve_sme_amd")):
    sys.path.insert(0, p)

import cgfdm  # noqa: E402  (descriptor structure only; no device call)
import finite_fault as ff  # noqa: E402
import pyoracle  # noqa: E402
import workloads  # noqa: E402

DH = 500.0


def selection():
    return np.array([r * 20 + c for r in range(0, 20, 2) for c in range(0, 20, 2)])


def main():
    w = workloads.finite_fault_workload(DH=DH)
    NX, NY, NZ = w["grid"]
    d = cgfdm.make_desc(NX, NY, NZ, DH=DH, DT=w["DT"], NT=w["NT"], nPML=w["nPML"], use_terrain=1, depth=w["depth"])
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "fault.bin")
        workloads.write_rupture(path, w["rupture"])
        xyz, table, tdt = ff.finite_fault_map(path, d, w["terrain"], w["lon0"], w["lat0"])
    print(f"{len(xyz)} source nodes, table nt {table.shape[0]}, dt {tdt}", flush=True)
    o = pyoracle.Oracle(d, pyoracle.F64)
    o.set_model(w["terrain"], *w["medium"])
    o.set_source_table(xyz, table, tdt)
    st = w["stations"]
    o.set_stations(st)
    t0 = time.time()
    for n in range(0, w["NT"], 50):
        o.step(min(50, w["NT"] - n))
        print(f"step {min(n + 50, w['NT'])}/{w['NT']}  {time.time() - t0:.0f} s", flush=True)
    sel = selection()
    out = os.path.join(HERE, "config5_reduced_fp64.npz")
    np.savez_compressed(out, stations=o.stations()[sel, :, :3].astype(np.float32), pgv=o.pgv().astype(np.float32),
                        station_xyz=st[sel], source_xyz=xyz, nt=w["NT"])
    print("wrote", out)


if __name__ == "__main__":
    main()
