"""Benchmark of the build-and-query path on one MI355X.

Workload (the 10M-point float3 case of the builder/query scaling config):
  * tree: 10M float3 points, uniform in [0,1)^3 (SplitMix64 seed 1),
    left-balanced round-robin tree built by cukd_build_tree;
  * step: one kNN (k=8, FixedCandidateList, stack-based traversal) batch
    of 1M uniform float3 queries (seed 2) over the resident tree.

Prints one JSON line.  value = queries/s of the kNN step; the build time,
FCP rate and the other headline figures ride along under "metrics".
Inputs are resident in HBM before any timed region starts.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def _pct(a, p):
    a = sorted(a)
    return a[min(len(a) - 1, int(round(p / 100.0 * (len(a) - 1))))]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--nq", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--build-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import cudakdtree_amd as ck
    from cudakdtree_amd import datagen

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    lay = ck.float_points(3)
    S = 12
    n, nq, k = args.n, args.nq, args.k

    pts_h = datagen.uniform(n, 3, 1)
    q_h = datagen.uniform(nq, 3, 2)
    pts = torch.from_numpy(pts_h).to(dev)
    queries = torch.from_numpy(q_h).to(dev)
    tree = torch.empty_like(pts)
    torch.cuda.synchronize()
    print("[bench] inputs resident: %d points, %d queries" % (n, nq), file=sys.stderr, flush=True)

    # ---------------- build: 10M float3 ----------------
    stream = torch.cuda.current_stream()
    for _ in range(2):
        tree.copy_(pts)
        ck.buildTree(tree, lay)
    torch.cuda.synchronize()
    build_ms = []
    for _ in range(args.build_steps):
        tree.copy_(pts)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ck.buildTree(tree, lay)
        e1.record(stream)
        torch.cuda.synchronize()
        build_ms.append(e0.elapsed_time(e1))
    L = int(math.floor(math.log2(n - 1))) + 1
    build_bytes = n * S + L * 2 * n * (S + 4)
    build_med = _pct(build_ms, 50)
    print("[bench] build 10M: median %.3f ms" % build_med, file=sys.stderr, flush=True)

    # ---------------- kNN step (the headline value) ----------------
    ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    d2 = torch.empty((nq, k), dtype=torch.float32, device=dev)
    ret = torch.empty(nq, dtype=torch.float32, device=dev)
    out = (ids, d2, ret)
    # plain launch order, for reference
    for _ in range(args.warmup):
        ck.knn(tree, lay, queries, k, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ck.knn(tree, lay, queries, k, out=out)
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # headline: Morton-ordered launch (ordering time included)
    flags = ck.QUERY_SORT
    for _ in range(args.warmup):
        ck.knn(tree, lay, queries, k, out=out, flags=flags)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        ck.knn(tree, lay, queries, k, out=out, flags=flags)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    qps = nq / (ms_per_step / 1e3)

    # node visits of this exact traversal (CUKD_STATS semantics)
    counter = torch.zeros(1, dtype=torch.int64, device=dev)
    ck.set_traversal_stats(counter)
    ck.knn(tree, lay, queries, k, out=out, flags=flags)
    torch.cuda.synchronize()
    ck.set_traversal_stats(None)
    visits = int(counter.item())
    knn_bytes = nq * (S + k * 8 + 4) + visits * S
    knn_gbs = knn_bytes / (ev_ms / 1e3) / 1e9
    print("[bench] knn k=%d: %.3f ms/step (events %.3f ms), visits %.1f/query"
          % (k, ms_per_step, ev_ms, visits / nq), file=sys.stderr, flush=True)

    # FCP rate on the same tree
    fids = torch.empty(nq, dtype=torch.int32, device=dev)
    fd2 = torch.empty(nq, dtype=torch.float32, device=dev)
    for _ in range(args.warmup):
        ck.fcp(tree, lay, queries, out=(fids, fd2), flags=flags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ck.fcp(tree, lay, queries, out=(fids, fd2), flags=flags)
    torch.cuda.synchronize()
    fcp_ms = (time.perf_counter() - t0) * 1e3 / args.steps

    # ---------------- CPU baseline (oracle, host cores) ----------------
    cpu = None
    if not args.no_cpu:
        try:
            import oracle as oc
            threads = min(16, os.cpu_count() or 1)
            tree_h = tree.cpu().numpy()
            qs = q_h[: args.cpu_sample]
            t0 = time.perf_counter()
            oc.knn(oc.STACK_BASED, oc.LIST_FIXED, k, tree_h, lay, qs, threads=threads)
            cpu_s = time.perf_counter() - t0
            cpu = {"value": round(len(qs) / cpu_s, 1), "unit": "queries/s", "cores": threads,
                   "kind": "port",
                   "sample": "%d of the %d kNN(k=%d) queries over the same 10M-point tree, "
                             "oracle stack-based traversal, OpenMP over queries"
                             % (len(qs), nq, k)}
            nb = 1_000_000
            t0 = time.perf_counter()
            oc.build_host(pts_h[:nb], lay)
            cpu["build_ms_1M_single_core"] = round((time.perf_counter() - t0) * 1e3, 1)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "queries/s", "cores": 0, "kind": "port",
                   "sample": "failed: %s" % e}

    line = {
        "metric": "build ms @10M float3; kNN(k=8) queries/s; batched-vs-plain speedup; p99 ms",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: uniform [0,1)^3 float3 points and queries (SplitMix64 seeds 1, 2)",
        "config": {"workload": "kNN k=8 (FixedCandidateList, stack-based, Morton-ordered "
                               "launch) of 1M float3 queries over a 10M-point left-balanced tree",
                   "points": n, "queries": nq, "k": k},
        "metrics": {
            "build_ms_10M_float3": round(build_med, 3),
            "build_ms_all": [round(x, 3) for x in build_ms],
            "build_roofline_frac": round(build_bytes / (build_med / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "knn_k8_queries_per_s": round(qps, 1),
            "knn_k8_queries_per_s_unordered": round(nq / (plain_ms / 1e3), 1),
            "fcp_queries_per_s": round(nq / (fcp_ms / 1e3), 1),
            "batched_vs_plain_speedup": None,
            "p99_ms": None,
            "visits_per_query": round(visits / nq, 2),
        },
        "roofline": {"bound": "hbm", "achieved": round(knn_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(knn_gbs / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "kernel": "k_query<3, stack-based, FixedCandidateList<8>> + Morton ordering",
                     "algorithmic_bytes_per_launch": knn_bytes,
                     "avg_launch_ms": round(ev_ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
