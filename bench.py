#!/usr/bin/env python3
"""Throughput benchmark of the CGFDM time loop on one MI355X.

Workload (the throughput configuration): 1001 x 1001 x 401 grid, dh = 100 m, DT = 5 ms,
3-layer medium (Vs 2000/3000/3500 m/s, interfaces 3 and 10 km below the surface, Vp = 1.73 Vs,
rho = 1740 (Vp/1000)^0.25) with a seeded 5 % random perturbation of ~1 km correlation, seeded
smooth random terrain of 1 km RMS, one double-couple source (strike 0, dip 90, rake 0,
M0 1e17, 2 Hz Ricker, 5 km deep) through the moment-rate table, 21 x 21 surface stations, PGV.
A step is one full time step (source, 4 fused RK stages, stations, PGV) over the whole grid;
inputs are resident in HBM before timing starts. GLUPS = NX*NY*NZ*steps/seconds.

The headline value is FP16-storage GLUPS (FP32 arithmetic); FP32-storage GLUPS is measured in
the same run and reported beside it, with the ratio. Prints one JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cgfdm-fp16_sve_sme_amd"))

import numpy as np  # noqa: E402

import cgfdm  # noqa: E402
import workloads as wl  # noqa: E402

METRIC = "GLUPS at 1001x1001x401 FP16-store vs FP32; batched p50/p99 solve latency"
PEAK_HBM_GBS = 8000.0
VALUES_PER_LUP = 192  # algorithmic values moved per grid point per full step (4 stages)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_model(NX, NY, NZ, DH, seed=2024):
    depth = NZ * DH
    terrain = wl.smooth_random_terrain(NX, NY, DH, rms=1000.0, corr=5000.0, seed=seed)
    vs, vp, rho = wl.build_medium(NX, NY, NZ, DH, terrain=terrain, depth=depth, pert=0.05,
                                  corr_m=1000.0, seed=seed + 1)
    return terrain, depth, vs, vp, rho


def source_and_stations(NX, NY, NZ, DH, DT, nt_src, rho_src, vs_src):
    mu = float(rho_src) * float(vs_src) ** 2
    table = wl.double_couple_table(1e17, 0.0, 90.0, 0.0, 2.0, nt_src, DT, mu)
    src = np.array([[NX // 2, NY // 2, NZ - 1 - int(round(5000.0 / DH))]], np.int32)
    st = wl.surface_stations(NX, NY, NZ, n=21)
    return src, table, st


def run_gpu(prec, grid, DH, DT, steps, warmup, model):
    NX, NY, NZ = grid
    terrain, depth, vs, vp, rho = model
    d = cgfdm.make_desc(NX, NY, NZ, DH=DH, DT=DT, NT=steps + warmup, nPML=12, precision=prec,
                        use_terrain=1, depth=depth)
    t0 = time.time()
    dom = cgfdm.Domain(d, device=0)
    dom.set_model(terrain, vs, vp, rho)
    src, table, st = source_and_stations(NX, NY, NZ, DH, DT, steps + warmup + 2,
                                         rho[src_k(NZ, DH), NY // 2, NX // 2],
                                         vs[src_k(NZ, DH), NY // 2, NX // 2])
    dom.set_source_table(src, table, DT)
    dom.set_stations(st)
    dom.sync()
    log(f"[{prec}] setup {time.time() - t0:.1f}s, device bytes {dom.bytes / 2**30:.1f} GiB")
    dom.step(warmup)
    dom.sync()
    dom.enable_timing(True)
    dom.stage_timing(reset=True)
    t0 = time.perf_counter()
    dom.step(steps)
    dom.sync()
    wall = time.perf_counter() - t0
    avg_ms, cnt = dom.stage_timing()
    sts = dom.stations()
    finite = bool(np.isfinite(sts).all())
    dom.close()
    n = NX * NY * NZ
    glups = n * steps / wall / 1e9
    log(f"[{prec}] {steps} steps in {wall:.3f}s: {glups:.3f} GLUPS, stage kernel avg {avg_ms:.3f} ms "
        f"over {cnt} launches, stations finite={finite}")
    return {"glups": glups, "ms_per_step": wall / steps * 1e3, "stage_avg_ms": avg_ms,
            "stage_launches": cnt, "finite": finite}


def src_k(NZ, DH):
    return NZ - 1 - int(round(5000.0 / DH))


def cpu_baseline(DH, DT, sample=(1001, 1001, 41), warmup=2, steps=5):
    """The FP32 CPU restatement of the reference scalar path (oracle), OpenMP over k-slabs on
    the host cores, timed on a z-slab sample of the same workload recipe."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.build()
    NX, NY, NZ = sample
    terrain, depth, vs, vp, rho = build_model(NX, NY, NZ, DH)
    d = cgfdm.make_desc(NX, NY, NZ, DH=DH, DT=DT, NT=steps + warmup, nPML=12, use_terrain=1,
                        depth=depth)
    o = pyoracle.Oracle(d, pyoracle.F32)
    o.set_model(terrain, vs, vp, rho)
    k = max(0, NZ - 1 - 10)
    src, table, st = source_and_stations(NX, NY, NZ, DH, DT, steps + warmup + 2,
                                         rho[k, NY // 2, NX // 2], vs[k, NY // 2, NX // 2])
    src[0, 2] = k
    o.set_source_table(src, table, DT)
    o.set_stations(st)
    o.step(warmup)
    t0 = time.perf_counter()
    o.step(steps)
    wall = time.perf_counter() - t0
    threads = pyoracle.lib().orc_num_threads()
    glups = NX * NY * NZ * steps / wall / 1e9
    log(f"[cpu] {NX}x{NY}x{NZ} x {steps} steps on {threads} threads: {wall:.2f}s, {glups:.4f} GLUPS")
    o.close()
    return {"value": glups, "unit": "GLUPS", "cores": threads, "kind": "port",
            "sample": f"{NX}x{NY}x{NZ} z-slab of the same recipe (terrain, layered+5% medium, "
                      f"double-couple table source, stations, PGV), FP32, {warmup} warm-up + "
                      f"{steps} timed steps, OpenMP over k on {threads} threads"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", default="1001,1001,401")
    ap.add_argument("--precisions", default="fp16,fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=5)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU slab "
              f"decomposition is not part of this build.", file=sys.stderr)
        sys.exit(2)
    grid = tuple(int(x) for x in args.grid.split(","))
    DH, DT = 100.0, 0.005
    NX, NY, NZ = grid
    t0 = time.time()
    model = build_model(NX, NY, NZ, DH)
    log(f"model {NX}x{NY}x{NZ} built in {time.time() - t0:.1f}s")
    res = {}
    for prec in args.precisions.split(","):
        res[prec] = run_gpu(prec, grid, DH, DT, args.steps, args.warmup, model)
    del model
    head = "fp16" if "fp16" in res else list(res)[0]
    r = res[head]
    sb = 4 if head == "fp32" else 2
    n = NX * NY * NZ
    bytes_per_launch = VALUES_PER_LUP / 4 * sb * n
    achieved = bytes_per_launch / (r["stage_avg_ms"] * 1e-3) / 1e9 if r["stage_avg_ms"] > 0 else None
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(DH, DT, steps=args.cpu_steps)
    out = {
        "metric": METRIC,
        "value": round(r["glups"], 4),
        "unit": "GLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"CGFDM time loop {NX}x{NY}x{NZ}, dh 100 m, DT 5 ms, random terrain "
                               f"1 km RMS, layered medium + 5% random, double-couple table source, "
                               f"21x21 stations, PGV, nPML 12, free surface",
                   "grid": [NX, NY, NZ], "storage": head, "arith": "f32"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4) if achieved else None,
                     "traffic": None,
                     "kernel": "fused RK stage (deriv + free surface + PML + RK update)",
                     "bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
    }
    for prec, v in res.items():
        out[f"{prec}_store"] = {"glups": round(v["glups"], 4), "ms_per_step": round(v["ms_per_step"], 3),
                                "stage_avg_ms": round(v["stage_avg_ms"], 4), "finite": v["finite"]}
    if "fp16" in res and "fp32" in res:
        out["fp16_over_fp32"] = round(res["fp16"]["glups"] / res["fp32"]["glups"], 3)
    out["latency_p50_ms"] = None
    out["latency_p99_ms"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
