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


def classify_traffic(n, nt):
    """HBM bytes per k_classify launch measured by the rocprofv3 FETCH_SIZE/WRITE_SIZE
    passes recorded in the committed summary (256^3 x 4 only; None otherwise)."""
    if (n, nt) != (256, 4) or not os.path.exists(TRAFFIC_SUMMARY):
        return None
    import re
    m = re.search(r"^classify_traffic_bytes:\s*(\d+)", open(TRAFFIC_SUMMARY).read(), re.M)
    return int(m.group(1)) if m else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(comps, n_sample):
    """The oracle (CPU restatement of the reference engine) on a bounded sample: the
    first n_sample^3 x 4 block of the same field, all host threads."""
    import oracle
    sub = [np.asfortranarray(c[:n_sample, :n_sample, :n_sample, :]) for c in comps]
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    threads = min(threads, 16)
    t0 = time.perf_counter()
    r = oracle.run(sub, oracle.PRED_CP, 3, nthreads=threads)
    dt = time.perf_counter() - t0
    n_simp = (n_sample - 1) ** 3 * (comps[0].shape[3] - 1) * N_TYPES
    return {"value": n_simp / dt, "unit": "simplices/s", "cores": threads, "kind": "port",
            "sample": f"{n_sample}^3x{comps[0].shape[3]} corner block of the same field "
                      f"({n_simp} tets, {len(r['verts'])} nodes, {dt:.2f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--nt", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 1 GiB device-to-device copy (PMC calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(multi-GPU sharding is designed, not built)"}))
        sys.exit(2)

    import torch
    from ftk2_amd import engine, synth

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
    classify_ms = float(ph[0])
    emit_ms = float(ph[4])
    dominant = "classify" if classify_ms >= emit_ms else "emit"
    if dominant == "classify":
        dom_bytes, dom_ms = in_bytes, classify_ms
    else:
        dom_bytes, dom_ms = n_nodes * (8 + 4 * 8) + n_edges * 8, emit_ms
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": classify_traffic(n, nt) if dominant == "classify" else None,
                     "algorithmic_bytes": int(dom_bytes)},
        "pipeline": {"kernel_ms": kernel_ms, "phase_ms": [float(x) for x in ph],
                     "phases": ["classify", "scan_nodes", "offsets+cell_count", "scan_cells+sync",
                                "emit", "link", "compress"],
                     "algorithmic_bytes": int(total_bytes), "input_bytes": int(in_bytes),
                     "hbm_frac_total": total_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "hbm_frac_input_only": in_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }
    if not args.no_cpu:
        log(f"[bench] cpu baseline on a {args.cpu_sample}^3 x {nt} sample")
        out["cpu_baseline"] = cpu_baseline(comps, args.cpu_sample)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
