"""Benchmark: 3D critical-point tracking, CriticalPointPredicate<3,double> on a
256x256x256 x 4-timestep RegularSimplicialMesh (FP64 gradient of 32 moving Gaussians +
1e-3 noise, 1% zero vertices, 1% duplicated values), one GPU.

A step is one ftk2h_extract_regular() call (execute() + get_complex() on device:
classify -> node offsets -> emit nodes/edges -> union-find) over inputs already
resident in HBM.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_TYPES = 60  # tetrahedra per 4-D hypercube
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "cp3d_256x4_fp64_hbm_traffic.md")


def kernel_source_hash():
    """SHA-256 over the library's kernel and host sources (ftk2_amd/csrc), recorded in
    the PMC traffic summary so that a summary of other kernels is never used."""
    import hashlib
    h = hashlib.sha256()
    d = os.path.join(ROOT, "ftk2_amd", "csrc")
    for name in sorted(os.listdir(d)):
        if name.endswith((".hip", ".hpp", ".cpp", ".h")):
            h.update(name.encode())
            h.update(open(os.path.join(d, name), "rb").read())
    return h.hexdigest()


def measured_traffic(kernel, n, nt):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, calibrated) from the
    committed rocprofv3 PMC summary of the 256^3 x 4 workload, or None when there is no
    summary for these sources (its recorded source hash differs)."""
    if (n, nt) != (256, 4) or not os.path.exists(TRAFFIC_SUMMARY):
        return None, "no PMC summary for this workload"
    import re
    txt = open(TRAFFIC_SUMMARY).read()
    m = re.search(r"^source_sha256:\s*([0-9a-f]{64})", txt, re.M)
    if not m or m.group(1) != kernel_source_hash():
        return None, "PMC summary made on other kernel sources"
    m = re.search(r"^" + re.escape(kernel) + r"_traffic_bytes:\s*(\d+)", txt, re.M)
    return (int(m.group(1)), "rocprofv3 PMC summary") if m else (None, "kernel not in the PMC summary")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": affinity}


def cpu_baseline(comps, n_sample, n_one):
    """The oracle (CPU restatement of the reference engine) on bounded samples of the
    same field: the first n_sample^3 x T block with the host's CPU share (all CPUs this
    process may use, at most 16: the GPU host's share per GPU), and the first
    n_one^3 x T block on one core."""
    import oracle
    info = host_info()
    nt = comps[0].shape[3]

    def timed(n, threads):
        sub = [np.asfortranarray(c[:n, :n, :n, :]) for c in comps]
        t0 = time.perf_counter()
        r = oracle.run(sub, oracle.PRED_CP, 3, nthreads=threads)
        dt = time.perf_counter() - t0
        n_simp = (n - 1) ** 3 * (nt - 1) * N_TYPES
        return n_simp / dt, f"{n}^3x{nt} corner block of the same field ({n_simp} tets, {len(r['verts'])} nodes, {dt:.2f} s)"

    threads = min(info["affinity"], 16)
    v_all, s_all = timed(n_sample, threads)
    v_one, s_one = timed(n_one, 1)
    return {"value": v_all, "unit": "simplices/s", "cores": threads, "kind": "port", "sample": s_all,
            "one_core": {"value": v_one, "cores": 1, "sample": s_one}, "host": info}


SCALAR_WORKLOADS = {
    # name: (predicate kind, m, simplex types per hypercube, side, timesteps, field seeds, threshold(s))
    "contour": (1, 1, 15, 384, 4, (0,), (0.0, 0.0)),
    "fiber": (2, 2, 50, 256, 4, (1, 2), (0.0, 0.0)),
}


def scalar_workload(args, name):
    """The contour (384^3 x 4 FP32, isovalue 0) and fiber (two 256^3 x 4 FP32 fields)
    workloads: fields = sum of 64 moving Gaussians minus the mean, inputs resident
    in HBM; same step and JSON schema as the main line."""
    import torch
    import oracle
    from ftk2_amd import engine, synth
    kind, m, ntypes, n, nt, seeds, thr = SCALAR_WORKLOADS[name]
    log(f"[bench] generating {len(seeds)} x {n}^3 x {nt} FP32 field(s)")
    comps = [synth.scalar_config(n=n, nt=nt, seed=s) for s in seeds]
    dev = [torch.from_numpy(np.ascontiguousarray(c.transpose())).cuda().permute(3, 2, 1, 0) for c in comps]
    torch.cuda.synchronize()
    mesh = engine.RegularSimplicialMesh([n, n, n, nt])
    for i in range(args.warmup):
        r = engine.extract(mesh, kind, m, dev, thresholds=thr)
        log(f"[bench] warmup {i}: phases(ms)={['%.3f' % x for x in r.phase_ms()]}")
        del r
    torch.cuda.synchronize()
    phases = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = engine.extract(mesh, kind, m, dev, thresholds=thr)
        phases.append(r.phase_ms())
        dv = r.device_view()
        n_nodes, n_cells = dv.n_nodes, (dv.conn_cells[0] if dv.n_conn else 0)
        del r
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ph = np.mean(np.array(phases), axis=0)
    log(f"[bench] {name}: ms/step={dt * 1e3:.3f} phases(ms)={['%.3f' % x for x in ph]} nodes={n_nodes} "
        f"cells={n_cells}")
    n_simp = (n - 1) ** 3 * (nt - 1) * ntypes
    in_bytes = len(comps) * n ** 3 * nt * 4
    # k_emit (phase 5) writes node key + lambda and the cells
    node_bytes = int(n_nodes) * (8 + (m + 1) * 4) + int(n_cells) * 4 * (3 if m == 2 else 4)
    achieved = node_bytes / (float(ph[5]) * 1e-3) / 1e9
    out = {
        "metric": f"simplices/s, {name} {n}^3x{nt} FP32",
        "value": n_simp / dt, "unit": "simplices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{name}_{n}^3x{nt}_fp32", "fields": len(comps), "gaussians": 64,
                   "thresholds": list(thr[:len(comps)]), "nodes": int(n_nodes), "cells": int(n_cells)},
        "roofline": {"bound": "hbm", "kernel": "k_emit", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes": int(node_bytes), "kernel_ms": float(ph[5])},
        "pipeline": {"phase_ms": [float(x) for x in ph]},
    }
    if not args.no_cpu:
        s = args.cpu_sample
        sub = [np.asfortranarray(c[:s, :s, :s, :]) for c in comps]
        threads = min(len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1, 16)
        log(f"[bench] cpu baseline on a {s}^3 x {nt} sample")
        t0 = time.perf_counter()
        ref = oracle.run(sub, {1: oracle.PRED_CONTOUR, 2: oracle.PRED_FIBER}[kind], m, thr=thr, nthreads=threads)
        cdt = time.perf_counter() - t0
        ns = (s - 1) ** 3 * (nt - 1) * ntypes
        out["cpu_baseline"] = {"value": ns / cdt, "unit": "simplices/s", "cores": threads, "kind": "port",
                               "sample": f"{s}^3x{nt} corner block of the same field(s) ({ns} simplices, "
                                         f"{len(ref['verts'])} nodes, {cdt:.2f} s)", "host": host_info()}
    return out


def batch_leg(seed, njobs, reps, log_every=256):
    """1024 independent CriticalPointPredicate<2,double> jobs (spatial sides log-uniform
    in [32, 1024], timesteps uniform in [2, 16], 4-64 moving Gaussians): per-job plain
    extraction vs one batched pass with the autotuned schedule."""
    import torch
    from ftk2_amd import _lib, engine, synth
    rng = np.random.default_rng(7919 + seed)
    jobs, simp, in_bytes = [], 0, 0
    t0 = time.perf_counter()
    for (nx, ny, nt), u, v in synth.batch_jobs(njobs, seed):
        jobs.append((engine.RegularSimplicialMesh([nx, ny, nt]), [u, v]))
        simp += (nx - 1) * (ny - 1) * (nt - 1) * 12
        in_bytes += 2 * nx * ny * nt * 8
    torch.cuda.synchronize()
    log(f"[batch] {njobs} jobs generated on the GPU in {time.perf_counter() - t0:.1f} s: {simp} triangles, "
        f"{in_bytes / 1e9:.2f} GB of fields")
    # plain: one ftk2h_extract_regular per job, sequentially
    engine.extract(jobs[0][0], _lib.PRED_CRITICAL_POINT, 2, jobs[0][1])  # warm
    t0 = time.perf_counter()
    n_nodes = 0
    for i, (mesh, comps) in enumerate(jobs):
        r = engine.extract(mesh, _lib.PRED_CRITICAL_POINT, 2, comps)
        n_nodes += r.sizes()[0]
        del r
        if (i + 1) % log_every == 0:
            log(f"[batch] plain {i + 1}/{njobs}")
    plain_s = time.perf_counter() - t0
    b = engine.Batch(_lib.PRED_CRITICAL_POINT, 2, _lib.F64, 3)
    t0 = time.perf_counter()
    cfg, tuned_ms = b.autotune(jobs, reps=1)
    log(f"[batch] autotune: {cfg} at {tuned_ms:.2f} ms ({time.perf_counter() - t0:.1f} s)")
    times = []
    for r_ in range(reps):
        t0 = time.perf_counter()
        res = b.run(jobs)
        times.append(time.perf_counter() - t0)
        if r_ < reps - 1:
            del res
    dev_ms = b.last_ms()
    # equal results: a sample of jobs against their plain extraction
    sample = rng.choice(njobs, size=min(8, njobs), replace=False)
    exact = True
    for i in sample:
        a = res[i].to_complex()
        p = engine.extract(jobs[i][0], _lib.PRED_CRITICAL_POINT, 2, jobs[i][1]).to_complex()
        exact = exact and np.array_equal(a.vertices, p.vertices) and np.array_equal(a.lam, p.lam) and \
            np.array_equal(a.track_id, p.track_id) and all(
                np.array_equal(a.connectivity[d], p.connectivity[d]) for d in p.connectivity)
    batch_nodes = sum(r.sizes()[0] for r in res)
    del res
    b.close()
    t_b = float(np.mean(times))
    return {"workload": "1024 x CriticalPointPredicate<2,double>, sides log-U[32,1024], T U[2,16], 4-64 Gaussians",
            "jobs": njobs, "simplices": int(simp), "input_bytes": int(in_bytes), "nodes": int(batch_nodes),
            "plain": {"simplices_per_s": simp / plain_s, "ms": plain_s * 1e3, "nodes": int(n_nodes)},
            "batched": {"simplices_per_s": simp / t_b, "ms": t_b * 1e3, "ms_runs": [x * 1e3 for x in times],
                        "device_ms": {"classify_to_bases": dev_ms[0], "emit_to_end": dev_ms[1]},
                        "config": cfg, "autotune_ms": tuned_ms},
            "speedup": plain_s / t_b, "bit_exact_sample": bool(exact) and batch_nodes == n_nodes}


def serve_leg(seed, n_req, rates_x, budget_gb):
    """Serving: Poisson arrivals of streaming-tracking requests, 70% 2D FP64 CP at
    128^2-2048^2 and 30% 3D FP64 CP at 64^3-256^3, 1-32 timesteps each, admitted
    against a device-memory budget and continuously batched.  Offered rate swept in
    multiples of the plain (one request at a time) service rate."""
    import torch
    from ftk2_amd import _lib, engine, serve, synth
    t0 = time.perf_counter()
    pools = synth.serving_pools(seed)
    torch.cuda.synchronize()
    log(f"[serve] request pools generated in {time.perf_counter() - t0:.1f} s")

    def make(nreq, rs):
        out = []
        for i, (sd, n, T) in enumerate(synth.serving_mix(nreq, rs)):
            pool = pools[(sd, n)]
            perm = tuple(reversed(range(sd)))
            steps = [[p[t].permute(*perm) for p in pool] for t in range(T)]
            simp = (n - 1) ** sd * max(0, T - 1) * (12 if sd == 2 else 60)
            out.append(serve.Request(i, 0.0, (n,) * sd, _lib.PRED_CRITICAL_POINT, sd, _lib.F64, steps,
                                     simplices=simp, meta={"sd": sd, "n": n, "T": T}))
        return out

    # one request set; the plain run and every offered rate serve exactly these
    reqs = make(n_req, np.random.default_rng(seed + 1))
    work = sum(r.simplices for r in reqs)

    def alone(rq):
        st = engine.Stream(rq.spatial, rq.kind, rq.m, rq.dtype, len(rq.timesteps[0]))
        for comps in rq.timesteps:
            st.append(comps)
        r = st.finish()
        st.close()
        return r

    alone(reqs[0])  # warm
    t0 = time.perf_counter()
    for rq in reqs:
        r = alone(rq)
        del r
    plain_s = time.perf_counter() - t0
    plain_rate = n_req / plain_s
    log(f"[serve] plain: {n_req} requests one at a time in {plain_s:.2f} s ({plain_rate:.1f} req/s, "
        f"{work / plain_s:.3e} simplices/s)")

    # each request alone in one extraction over all its timesteps (needs them all up
    # front, which a stream does not): the fastest single-request path
    def oneshot(rq):
        sd, n, T = rq.meta["sd"], rq.meta["n"], rq.meta["T"]
        if T < 2:
            return None
        pool = pools[(sd, n)]
        views = [p[:T].permute(*reversed(range(sd + 1))) for p in pool]
        return engine.extract(engine.RegularSimplicialMesh([n] * sd + [T]), _lib.PRED_CRITICAL_POINT, sd, views)

    oneshot(reqs[0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for rq in reqs:
        r = oneshot(rq)
        del r
    torch.cuda.synchronize()
    one_s = time.perf_counter() - t0
    log(f"[serve] one-shot: {n_req} requests one at a time in {one_s:.2f} s ({n_req / one_s:.1f} req/s, "
        f"{work / one_s:.3e} simplices/s)")
    points = []
    for k, x in enumerate(rates_x):
        rate = x * plain_rate
        arr = serve.poisson_arrivals(n_req, rate, np.random.default_rng(seed * 1000 + k))
        for rq, a in zip(reqs, arr):
            rq.arrival = float(a)
        srv = engine.Server(int(budget_gb * (1 << 30)))
        outs, wall, rej = serve.run(srv, reqs)
        pt = serve.summarize(outs, reqs, wall)
        pt.update({"offered_req_per_s": rate, "offered_x_plain": x, "rejected": len(rej)})
        points.append(pt)
        log(f"[serve] offered {rate:.1f} req/s ({x}x plain): p50 {pt['p50_ms']:.1f} ms p99 {pt['p99_ms']:.1f} ms "
            f"{pt['requests_per_s']:.1f} req/s {pt['throughput_simplices_per_s']:.3e} simplices/s")
        srv.close()
    # each result the same alone and in the mix (requests up to 512^2 / 128^3)
    check = [r for r in reqs if (r.meta["sd"] == 2 and r.meta["n"] <= 512) or
             (r.meta["sd"] == 3 and r.meta["n"] <= 128)][:8]
    for rq in check:
        rq.arrival = 0.0
    srv = engine.Server(int(budget_gb * (1 << 30)))
    outs, _, _ = serve.run(srv, check, keep_results=True)
    exact = len(outs) == len(check)
    by_rid = {r.rid: r for r in check}
    for o in outs:
        a, b = o.result.to_complex(), alone(by_rid[o.rid]).to_complex()
        exact = exact and np.array_equal(a.vertices, b.vertices) and np.array_equal(a.lam, b.lam) and \
            np.array_equal(a.track_id, b.track_id) and all(
                np.array_equal(a.connectivity[d], b.connectivity[d]) for d in b.connectivity)
    srv.close()
    sat = max(points, key=lambda p: p["throughput_simplices_per_s"])
    plain_tp = work / plain_s
    return {"workload": "Poisson streams: 70% 2D FP64 CP 128^2-2048^2, 30% 3D FP64 CP 64^3-256^3, 1-32 timesteps",
            "budget_gb": budget_gb, "requests_per_point": n_req,
            "plain": {"simplices_per_s": plain_tp, "requests_per_s": plain_rate, "wall_s": plain_s,
                      "how": "each request alone through ftk2h_stream, one slab per appended timestep"},
            "oneshot": {"simplices_per_s": work / one_s, "requests_per_s": n_req / one_s, "wall_s": one_s,
                        "how": "each request alone in one ftk2h_extract_regular over all its timesteps"},
            "points": points, "p50_ms": sat["p50_ms"], "p99_ms": sat["p99_ms"],
            "p99_ms_at_2x": next((p["p99_ms"] for p in points if p["offered_x_plain"] == 2), None),
            "saturated_simplices_per_s": sat["throughput_simplices_per_s"],
            "speedup_vs_plain": sat["throughput_simplices_per_s"] / plain_tp,
            "speedup_vs_oneshot": sat["throughput_simplices_per_s"] / (work / one_s),
            "requests_per_s_vs_plain": sat["requests_per_s"] / plain_rate,
            "alone_equals_mixed": bool(exact)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nt", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--cpu-sample-one", type=int, default=32, help="side of the one-core CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 1024-job batching leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--batch-jobs", type=int, default=1024)
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (PMC calibration)")
    ap.add_argument("--workload", choices=["cp3d", "contour", "fiber"], default="cp3d",
                    help="cp3d (the benchmark line), or the contour / fiber workloads on their own")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is designed, not built)"}))
        sys.exit(2)

    import torch
    from ftk2_amd import engine, synth

    if args.workload != "cp3d":
        print(json.dumps(scalar_workload(args, args.workload)), flush=True)
        return
    n, nt = args.n, args.nt
    log(f"[bench] generating {n}^3 x {nt} FP64 field (seed {args.seed})")
    t0 = time.perf_counter()
    comps = synth.cp3d_config(n=n, nt=nt, seed=args.seed)
    log(f"[bench] generated in {time.perf_counter() - t0:.1f} s")
    dev = [torch.from_numpy(np.ascontiguousarray(c.transpose())).cuda() for c in comps]
    # index [x, y, z, t] views of the device arrays (x fastest in memory)
    dev_views = [d.permute(3, 2, 1, 0) for d in dev]
    torch.cuda.synchronize()
    if args.calib_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        dst.copy_(src)  # reads and writes exactly 1 GiB each
        torch.cuda.synchronize()
        del src, dst
    mesh = engine.RegularSimplicialMesh([n, n, n, nt])

    def step():
        return engine.extract(mesh, 0, 3, dev_views)

    for i in range(args.warmup):
        r = step()
        log(f"[bench] warmup {i}: phases(ms)={['%.3f' % x for x in r.phase_ms()]}")
        del r
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phases = []
    n_nodes = n_edges = 0
    for i in range(args.steps):
        r = step()
        phases.append(r.phase_ms())
        dv = r.device_view()
        n_nodes = dv.n_nodes
        n_edges = dv.conn_cells[0] if dv.n_conn else 0
        del r
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ph = np.mean(np.array(phases), axis=0)
    log(f"[bench] ms/step={dt * 1e3:.3f} phases(ms)={['%.3f' % x for x in ph]} nodes={n_nodes} "
        f"edges={n_edges}")

    n_hc = (n - 1) ** 3 * (nt - 1)
    n_simp = n_hc * N_TYPES
    in_bytes = 3 * n ** 3 * nt * 8
    # algorithmic bytes of the whole path: inputs once + node key and lambda + edges
    # (two u32) + track id (u64), as the reference's record layout would count it
    total_bytes = in_bytes + n_nodes * (8 + 4 * 8) + n_edges * 8 + n_nodes * 8
    kernel_ms = float(ph.sum())
    # the largest single kernel (HIP events around each kernel on its stream) and its
    # algorithmic bytes: k_classify reads the inputs once; k_emit writes node key +
    # lambda and the edges; k_compress writes the track ids; stage 2 and the union-find
    # hooking have no algorithmic bytes of their own (they re-read what the first stage
    # decided) and are never named
    kernels = {"k_classify": (float(ph[0]), in_bytes),
               "k_stage2": (float(ph[1]), None),
               "k_emit": (float(ph[5]), n_nodes * (8 + 4 * 8) + n_edges * 8),
               "k_link_pairs": (float(ph[6]), None),
               "k_compress": (float(ph[7]), n_nodes * 8)}
    dominant = max((k for k in kernels if kernels[k][1] is not None), key=lambda k: kernels[k][0])
    dom_ms, dom_bytes = kernels[dominant]
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(dominant, n, nt)
    out = {
        "metric": "simplices/s + %HBM roofline, 3D CP 256^3x4 FP64; p50/p99 latency in serve loop",
        "value": n_simp / dt,
        "unit": "simplices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"cp3d_{n}^3x{nt}_fp64", "predicate": "CriticalPointPredicate<3,double>",
                   "mesh": [n, n, n, nt], "gaussians": 32, "noise": 1e-3, "zero_frac": 0.01,
                   "dup_frac": 0.01, "seed": args.seed, "nodes": int(n_nodes), "edges": int(n_edges)},
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "algorithmic_bytes": int(dom_bytes),
                     "kernel_ms": dom_ms},
        "pipeline": {"kernel_ms": kernel_ms, "phase_ms": [float(x) for x in ph],
                     "phases": ["k_classify", "k_stage2", "scan_nodes", "k_offsets", "scan_cells+sync",
                                "k_emit", "k_link_pairs", "k_compress"],
                     "algorithmic_bytes": int(total_bytes), "input_bytes": int(in_bytes),
                     "hbm_frac_total": total_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "hbm_frac_input_only": in_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    del dev, dev_views
    torch.cuda.empty_cache()
    if not args.no_batch:
        out["batching"] = batch_leg(args.seed, args.batch_jobs, reps=3)
        torch.cuda.empty_cache()
    if not args.no_serve:
        out["serving"] = serve_leg(args.seed, args.serve_requests, [0.5, 1, 2, 4, 8], args.serve_budget_gb)
        out["p50_ms"] = out["serving"]["p50_ms"]
        out["p99_ms"] = out["serving"]["p99_ms"]
        torch.cuda.empty_cache()
    if not args.no_cpu:
        log(f"[bench] cpu baseline on a {args.cpu_sample}^3 x {nt} sample and a {args.cpu_sample_one}^3 x {nt} "
            f"one-core sample")
        out["cpu_baseline"] = cpu_baseline(comps, args.cpu_sample, args.cpu_sample_one)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
