#!/usr/bin/env python3
"""Throughput benchmark of the CGFDM time loop on one MI355X.

Workload (the throughput configuration): 1001 x 1001 x 401 grid, dh = 100 m, DT = 5 ms,
3-layer medium (Vs 2000/3000/3500 m/s, interfaces 3 and 10 km below the surface, Vp = 1.73 Vs,
rho = 1740 (Vp/1000)^0.25) with a seeded 5 % random perturbation of ~1 km correlation, seeded
smooth random terrain of 1 km RMS, one double-couple source (strike 0, dip 90, rake 0,
M0 1e16, 2 Hz Ricker, 5 km deep) through the moment-rate table, 21 x 21 surface stations, PGV.
The moment keeps the near-source field inside the FP16 storage range at the reference's scales
(Cv = 5e3: |v| < 13 m/s): on this recipe 1e16 N m peaks at 2.1 m/s, 1e17 N m at 21 m/s, which
overflows FP16 storage after ~150 steps.
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
# values the bulk kernel itself moves per point and step: the four metric components that are
# exactly zero on terrain-following grids (xi_y, xi_z, eta_x, eta_z) are not read (4 x 4 fewer),
# zeta_x, zeta_y, zeta_z come from per-column values (3 x 4 fewer), and the RK accumulator after
# stage 0 is rebuilt in stage 1 instead of being written and read back (9 + 9 fewer); when xi_x
# and eta_y are the same at every point (terrain-following grids) they come from one stored copy
# (2 x 4 fewer)
VALUES_PER_LUP_MOVED = 146
VALUES_PER_LUP_MOVED_UNIFORM_PAIR = 138


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
    table = wl.double_couple_table(1e16, 0.0, 90.0, 0.0, 2.0, nt_src, DT, mu)
    src = np.array([[NX // 2, NY // 2, NZ - 1 - int(round(5000.0 / DH))]], np.int32)
    st = wl.surface_stations(NX, NY, NZ, n=21)
    return src, table, st


def run_gpu(prec, grid, DH, DT, steps, warmup, model):
    NX, NY, NZ = grid
    terrain, depth, vs, vp, rho = model
    NT = warmup + steps + min(steps, 4)
    d = cgfdm.make_desc(NX, NY, NZ, DH=DH, DT=DT, NT=NT, nPML=12, precision=prec,
                        use_terrain=1, depth=depth)
    t0 = time.time()
    dom = cgfdm.Domain(d, device=0)
    dom.set_model(terrain, vs, vp, rho)
    src, table, st = source_and_stations(NX, NY, NZ, DH, DT, NT + 2,
                                         rho[src_k(NZ, DH), NY // 2, NX // 2],
                                         vs[src_k(NZ, DH), NY // 2, NX // 2])
    dom.set_source_table(src, table, DT)
    dom.set_stations(st)
    dom.sync()
    log(f"[{prec}] setup {time.time() - t0:.1f}s, device bytes {dom.bytes / 2**30:.1f} GiB")
    dom.step(warmup)
    dom.sync()
    # timed region: `steps` whole steps, no per-kernel events (they would add host work)
    t0 = time.perf_counter()
    dom.step(steps)
    dom.sync()
    wall = time.perf_counter() - t0
    # kernel split over the same number of further steps, with HIP events on the domain's stream
    dom.enable_timing(True)
    dom.kernel_timing(reset=True)
    dom.step(min(steps, 4))
    (stage_ms, bulk_ms, gen_ms), cnt = dom.kernel_timing()
    dom.enable_timing(False)
    bulk_pts, gen_pts, tile, kc = dom.partition()
    uniform_pair = dom.uniform_metric_pair()
    sts = dom.stations()
    finite = bool(np.isfinite(sts).all())
    dom.close()
    n = NX * NY * NZ
    glups = n * steps / wall / 1e9
    log(f"[{prec}] {steps} steps in {wall:.3f}s: {glups:.3f} GLUPS; per stage {stage_ms:.3f} ms = bulk "
        f"{bulk_ms:.3f} ms ({bulk_pts / 1e6:.1f} M pts, tile {tile}, kc {kc}) + boundary {gen_ms:.3f} ms "
        f"({gen_pts / 1e6:.1f} M pts), stations finite={finite}")
    return {"glups": glups, "ms_per_step": wall / steps * 1e3, "stage_avg_ms": stage_ms,
            "bulk_ms": bulk_ms, "boundary_ms": gen_ms, "bulk_points": bulk_pts, "boundary_points": gen_pts,
            "tile_cfg": tile, "kc": kc, "uniform_metric_pair": uniform_pair, "stage_launches": cnt,
            "finite": finite}


def src_k(NZ, DH):
    return NZ - 1 - int(round(5000.0 / DH))


def cpu_baseline(DH, DT, sample=(1001, 1001, 41), warmup=2, steps=10):
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
            "per_core_mlups": round(glups * 1e3 / threads, 4),
            "reference_per_core_mlups": 0.72,
            "sample": f"{NX}x{NY}x{NZ} z-slab of the same recipe (terrain, layered+5% medium, "
                      f"double-couple table source, stations, PGV), FP32, {warmup} warm-up + "
                      f"{steps} timed steps, OpenMP over k on {threads} threads; the reference's own "
                      f"FP32-CGFDM rate is 0.72 MLUPS per ARM core (64 ranks, 601x601x151 Tibet run)"}


def serving_run(reqs, max_domains, steps_per_launch, paced=True, use_graph=True, least_work_first=False):
    """Serves `reqs` [(arrival_s, request)] with one cgfdm.Server; requests are submitted at
    their arrival times by a second thread (paced) or all at once. Returns latencies (s, from
    submission to results on the host), the makespan and the server counters."""
    import threading
    descs = []
    for _, r in reqs:
        NX, NY, NZ = r["grid"]
        descs.append(cgfdm.make_desc(NX, NY, NZ, DH=r["DH"], DT=r["DT"], NT=r["NT"], nPML=r["nPML"],
                                     precision=r["precision"]))
    srv = cgfdm.Server(precision=reqs[0][1]["precision"], max_domains=max_domains,
                       steps_per_launch=steps_per_launch, use_graph=use_graph, least_work_first=least_work_first)
    ids = {}
    err = []

    def feeder():
        t0 = time.perf_counter()
        for (t, r), d in zip(reqs, descs):
            if paced:
                dt = t - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
            try:
                ids[srv.submit({"desc": d, "layers": r["layers"], "table": r["table"],
                                "stations": r["stations"]})] = d
            except Exception as e:  # reported by the main thread
                err.append(e)
                return

    th = threading.Thread(target=feeder)
    t0 = time.perf_counter()
    th.start()
    lat, queue, service, first, last = [], [], [], None, 0.0
    done = 0
    while done < len(reqs):
        if err:
            raise err[0]
        got = srv.wait(timeout_ms=120000)
        if not got:
            raise RuntimeError(f"serving stalled: {srv.stats()}")
        for rid in got:
            rc, st, pgv, info = srv.result(rid)
            if rc != 0:
                raise RuntimeError(f"request {rid} failed: {info['error']}")
            if not np.isfinite(st).all():
                raise RuntimeError(f"request {rid}: non-finite seismograms")
            lat.append(info["t_done"] - info["t_submit"])
            queue.append(info["t_admit"] - info["t_submit"])
            service.append(info["t_done"] - info["t_admit"])
            first = info["t_submit"] if first is None else min(first, info["t_submit"])
            last = max(last, info["t_done"])
            done += 1
            if done % 32 == 0 and done < len(reqs):
                log(f"  {done}/{len(reqs)} requests served, {time.perf_counter() - t0:.0f}s")
    th.join()
    stats = srv.stats()
    srv.close()
    wall = time.perf_counter() - t0
    return np.array(lat), last - first, wall, stats, np.array(queue), np.array(service)


def serving(n, seed, max_domains, steps_per_launch, use_graph=True, modes=("batched", "one_at_a_time"),
            least_work_first=False, rate_hz=20.0, sustained_hz=0.0):
    reqs = wl.serving_stream(n, seed, rate_hz=rate_hz)
    lups = sum(r["grid"][0] * r["grid"][1] * r["grid"][2] * r["NT"] for _, r in reqs)
    out = {}
    for name, md in (("batched", max_domains), ("one_at_a_time", 1)):
        if name not in modes:
            continue
        lat, span, wall, st, qt, sv = serving_run(reqs, md, steps_per_launch, use_graph=use_graph,
                                                  least_work_first=least_work_first)
        out[name] = {"solves_per_s": round(n / span, 4), "p50_s": round(float(np.percentile(lat, 50)), 3),
                     "p99_s": round(float(np.percentile(lat, 99)), 3), "makespan_s": round(span, 2),
                     "glups": round(lups / span / 1e9, 3), "peak_resident": st["peak_resident"],
                     "launches": st["launches"], "rescaled": st["rescaled"],
                     "queue_p50_s": round(float(np.percentile(qt, 50)), 3),
                     "queue_p99_s": round(float(np.percentile(qt, 99)), 3),
                     "service_p50_s": round(float(np.percentile(sv, 50)), 3),
                     "service_p99_s": round(float(np.percentile(sv, 99)), 3)}
        log(f"[serving {name}] {n} requests in {span:.1f}s: {n / span:.3f} solves/s, p50 "
            f"{np.percentile(lat, 50):.2f}s p99 {np.percentile(lat, 99):.2f}s (queued p50 {np.percentile(qt, 50):.2f}s, "
            f"service p50 {np.percentile(sv, 50):.2f}s p99 {np.percentile(sv, 99):.2f}s), {lups / span / 1e9:.2f} GLUPS, "
            f"peak resident {st['peak_resident']}")
    if sustained_hz > 0 and "batched" in modes:
        # the same requests arriving at a rate the GPU sustains: latency is then mostly service
        slow = wl.serving_stream(n, seed, rate_hz=sustained_hz)
        lat, span, wall, st, qt, sv = serving_run(slow, max_domains, steps_per_launch, use_graph=use_graph,
                                                  least_work_first=least_work_first)
        out["batched_sustained"] = {"arrival_hz": sustained_hz, "solves_per_s": round(n / span, 4),
                                    "p50_s": round(float(np.percentile(lat, 50)), 3),
                                    "p99_s": round(float(np.percentile(lat, 99)), 3),
                                    "makespan_s": round(span, 2), "peak_resident": st["peak_resident"],
                                    "queue_p50_s": round(float(np.percentile(qt, 50)), 3),
                                    "service_p50_s": round(float(np.percentile(sv, 50)), 3),
                                    "service_p99_s": round(float(np.percentile(sv, 99)), 3)}
        log(f"[serving batched, arrivals at {sustained_hz}/s] {n} requests in {span:.1f}s: p50 "
            f"{np.percentile(lat, 50):.2f}s p99 {np.percentile(lat, 99):.2f}s (service p50 {np.percentile(sv, 50):.2f}s), "
            f"peak resident {st['peak_resident']}")
    out["requests"] = n
    out["admit_order"] = "least work first" if least_work_first else "arrival"
    out["stream"] = (f"first {n} of the seeded 512-request stream (seed {seed}): Poisson arrivals at {rate_hz:g}/s, "
                     f"grids U[101,301]^2 x U[51,151], 2-5 random layers Vs 1500-4000 m/s, random double "
                     f"couples (M0 1e15 N m) 1-15 km deep, Ricker 0.5-3 Hz, 500-1500 steps, FP16 with scales chosen per request from the source, 5x5 surface "
                     f"stations")
    if "batched" in out and "one_at_a_time" in out:
        out["speedup"] = round(out["batched"]["solves_per_s"] / out["one_at_a_time"]["solves_per_s"], 3)
    return out


def finite_fault_run(steps, prec="fp16", DH=100.0):
    """The kinematic finite-fault workload (workloads.finite_fault_workload: 1201 x 1201 x 401,
    80 x 20 subfaults written as a reference-layout source file and mapped by
    cgfdm_finite_fault_map, fractal terrain, perturbed layered medium, 400 stations, PGV), FP16
    storage with scales chosen from the source; `steps` steps (4000 = the whole 32 s), progress
    every 500. Returns GLUPS, ms per step and the peak ground velocity."""
    import tempfile
    t0 = time.time()
    w = wl.finite_fault_workload(DH=DH)
    NX, NY, NZ = w["grid"]
    log(f"[finite fault] model {NX}x{NY}x{NZ} built in {time.time() - t0:.1f}s")
    d32 = cgfdm.make_desc(NX, NY, NZ, DH=DH, DT=w["DT"], NT=steps, nPML=w["nPML"], use_terrain=1, depth=w["depth"])
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "fault.bin")
        wl.write_rupture(path, w["rupture"])
        xyz, table, tdt = cgfdm.finite_fault_map(path, d32, w["terrain"], w["lon0"], w["lat0"])
    vs, vp, rho = w["medium"]
    sc = {}
    if prec != "fp32":
        Cv, Cs = cgfdm.source_scales(DH, xyz, table, tdt, float((rho.astype(np.float64) * vs * vs).max()))
        sc = {"Cv": Cv, "Cs": Cs, "Crho": 1e6}
    d = cgfdm.make_desc(NX, NY, NZ, DH=DH, DT=w["DT"], NT=steps, nPML=w["nPML"], precision=prec, use_terrain=1,
                        depth=w["depth"], check_cfl=1, **sc)
    dom = cgfdm.Domain(d, device=0)
    dom.set_model(w["terrain"], vs, vp, rho)
    dom.set_source_table(xyz, table, tdt)
    dom.set_stations(w["stations"])
    dom.sync()
    log(f"[finite fault] {len(xyz)} source nodes, scales {sc}, device bytes {dom.bytes / 2**30:.1f} GiB")
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(500, steps - done)
        dom.step(n)
        dom.sync()
        done += n
        log(f"[finite fault] {done}/{steps} steps, {time.perf_counter() - t0:.1f}s")
    wall = time.perf_counter() - t0
    pgv = dom.pgv()
    st = dom.stations()
    dom.close()
    glups = NX * NY * NZ * steps / wall / 1e9
    out = {"glups": round(glups, 4), "ms_per_step": round(wall / steps * 1e3, 3), "steps": steps,
           "grid": [NX, NY, NZ], "storage": prec, "source_nodes": int(len(xyz)),
           "pgv_max_m_s": round(float(pgv[:, :, 1].max()), 4), "finite": bool(np.isfinite(st).all() and
                                                                               np.isfinite(pgv).all())}
    log(f"[finite fault] {steps} steps in {wall:.1f}s: {glups:.3f} GLUPS, PGV max {out['pgv_max_m_s']} m/s, "
        f"finite={out['finite']}")
    return out


def copy_probe(gib):
    """A plain device copy of known size, for calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE."""
    import torch
    n = int(gib * 2**30)
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"copy probe: 3 copies of {n} bytes")


def measured_traffic(prec, grid):
    """HBM bytes per bulk-kernel launch from the committed counter summary of this workload, if
    one exists (profiles/traffic_*.json, written from two rocprofv3 --pmc runs)."""
    fn = os.path.join(ROOT, "profiles", f"traffic_bulk_{prec}_{grid[0]}x{grid[1]}x{grid[2]}.json")
    if not os.path.exists(fn):
        return None
    with open(fn) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", default="1001,1001,401")
    ap.add_argument("--precisions", default="fp16,fp32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--serve-requests", type=int, default=24,
                    help="requests of the serving stream to run (0: skip; 512: the whole stream)")
    ap.add_argument("--serve-slots", type=int, default=32)
    ap.add_argument("--serve-steps-per-launch", type=int, default=8)
    ap.add_argument("--serve-graph", type=int, default=1, help="1: replay captured HIP graphs, 0: launch directly")
    ap.add_argument("--serve-modes", default="batched,one_at_a_time", help="comma list of batched, one_at_a_time")
    ap.add_argument("--serve-order", default="arrival", choices=("arrival", "least-work"),
                    help="which waiting request a free slot takes")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrival rate of the stream (1/s)")
    ap.add_argument("--serve-sustained-rate", type=float, default=1.0,
                    help="also serve the same requests batched at this arrival rate, one the GPU sustains "
                         "(0: skip); reported as serving.batched_sustained")
    ap.add_argument("--finite-fault-steps", type=int, default=0,
                    help="also run the kinematic finite-fault workload (1201x1201x401, FP16) for this many "
                         "steps (4000: the whole rupture) and report it under 'finite_fault'")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run a device copy of this many GiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU slab "
              f"decomposition is not part of this build.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    grid = tuple(int(x) for x in args.grid.split(","))
    DH, DT = 100.0, 0.005
    NX, NY, NZ = grid
    res = {}
    precs = [p for p in args.precisions.split(",") if p]
    if precs:
        t0 = time.time()
        model = build_model(NX, NY, NZ, DH)
        log(f"model {NX}x{NY}x{NZ} built in {time.time() - t0:.1f}s")
        for prec in precs:
            res[prec] = run_gpu(prec, grid, DH, DT, args.steps, args.warmup, model)
        del model
    ffault = finite_fault_run(args.finite_fault_steps) if args.finite_fault_steps > 0 else None
    serve = serving(args.serve_requests, 7, args.serve_slots, args.serve_steps_per_launch, bool(args.serve_graph),
                    tuple(args.serve_modes.split(",")), args.serve_order == "least-work", args.serve_rate,
                    args.serve_sustained_rate) if args.serve_requests > 0 else None
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(DH, DT, steps=args.cpu_steps)
    head = "fp16" if "fp16" in res else (list(res)[0] if res else None)
    out = {"metric": METRIC, "unit": "GLUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic"}
    if head:
        r = res[head]
        sb = 4 if head == "fp32" else 2
        # dominant kernel: the bulk stage kernel; 48 values per point per launch on average
        bulk_bytes = VALUES_PER_LUP / 4 * sb * r["bulk_points"]
        achieved = bulk_bytes / (r["bulk_ms"] * 1e-3) / 1e9 if r["bulk_ms"] > 0 else None
        tr = measured_traffic(head, grid)
        moved = VALUES_PER_LUP_MOVED_UNIFORM_PAIR if r["uniform_metric_pair"] else VALUES_PER_LUP_MOVED
        out.update({
            "value": round(r["glups"], 4),
            "ms_per_step": round(r["ms_per_step"], 3),
            "config": {"workload": f"CGFDM time loop {NX}x{NY}x{NZ}, dh 100 m, DT 5 ms, random terrain "
                                   f"1 km RMS, layered medium + 5% random, double-couple table source M0 1e16 N m, "
                                   f"21x21 stations, PGV, nPML 12, free surface",
                       "grid": [NX, NY, NZ], "storage": head, "arith": "f32"},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
                         "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(achieved / PEAK_HBM_GBS, 4) if achieved else None,
                         "traffic": tr["bytes_per_launch"] if tr else None,
                         "kernel": "k_stage_bulk (fused RK stage over the bulk box)",
                         "bytes_per_launch": bulk_bytes, "launch_ms": round(r["bulk_ms"], 4),
                         "traffic_source": tr["source"] if tr else None,
                         "own_count": {"values_per_lup": moved,
                                       "achieved": round(achieved * moved / VALUES_PER_LUP, 1)
                                       if achieved else None,
                                       "frac": round(achieved * moved / VALUES_PER_LUP
                                                     / PEAK_HBM_GBS, 4) if achieved else None},
                         "whole_stage": {"ms": round(r["stage_avg_ms"], 4),
                                         "achieved": round(VALUES_PER_LUP / 4 * sb * NX * NY * NZ
                                                           / (r["stage_avg_ms"] * 1e-3) / 1e9, 1)}},
        })
    out["cpu_baseline"] = cpu
    for prec, v in res.items():
        out[f"{prec}_store"] = {k: (round(x, 4) if isinstance(x, float) else x) for k, x in v.items()}
    if "fp16" in res and "fp32" in res:
        out["fp16_over_fp32"] = round(res["fp16"]["glups"] / res["fp32"]["glups"], 3)
    if ffault:
        out["finite_fault"] = ffault
    if serve and "batched" in serve:
        out["latency_p50_ms"] = round(serve["batched"]["p50_s"] * 1e3, 1)
        out["latency_p99_ms"] = round(serve["batched"]["p99_s"] * 1e3, 1)
        out["serving"] = serve
    else:
        out["latency_p50_ms"] = None
        out["latency_p99_ms"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
