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


def forest_bench(args, ck, datagen, torch, np):
    """Batched vs plain over a forest of independent trees: 4096 float3
    trees with sizes log-uniform in [64, 262144] (about 129M points), each
    uniform in its own random axis-aligned box; per tree 0.5x its size kNN
    (k=4) queries uniform in the same box.  Plain = one cukd_build_tree and
    one cukd_knn per tree; batched = cukd_forest_build + cukd_forest_knn,
    captured once into a HIP graph and replayed.  Both give identical trees
    and results (checked here)."""
    dev = torch.device("cuda:0")
    T, k = args.forest_trees, 4
    r = (datagen.splitmix64(41, T) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    sizes = np.exp(np.log(64) + r * (np.log(262144) - np.log(64))).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    nq_t = sizes // 2
    qoff = np.concatenate([[0], np.cumsum(nq_t)]).astype(np.int64)
    N, NQ = int(off[-1]), int(qoff[-1])
    g = torch.Generator(device=dev)
    g.manual_seed(43)
    lo = torch.rand((T, 3), generator=g, device=dev) * 100.0
    ext = torch.rand((T, 3), generator=g, device=dev) * 10.0 + 0.1
    tid = torch.repeat_interleave(torch.arange(T, device=dev), torch.from_numpy(sizes).to(dev))
    qtid = torch.repeat_interleave(torch.arange(T, device=dev), torch.from_numpy(nq_t).to(dev))
    orig = (torch.rand((N, 3), generator=g, device=dev) * ext[tid] + lo[tid]).contiguous()
    queries = (torch.rand((NQ, 3), generator=g, device=dev) * ext[qtid] + lo[qtid]).contiguous()
    del tid, qtid
    lay = ck.float_points(3)
    seg_off = torch.from_numpy(qoff.astype(np.int32)).to(dev)
    seg_tree = torch.arange(T, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    print("[bench] forest: %d trees, %d points, %d queries" % (T, N, NQ), file=sys.stderr,
          flush=True)

    # plain: one build + one kNN launch sequence per tree
    p_pts = torch.empty_like(orig)
    p_ids = torch.empty((NQ, k), dtype=torch.int32, device=dev)
    p_d2 = torch.empty((NQ, k), dtype=torch.float32, device=dev)
    p_r = torch.empty(NQ, dtype=torch.float32, device=dev)

    def plain():
        p_pts.copy_(orig)
        for t in range(T):
            a, b, qa, qb = int(off[t]), int(off[t + 1]), int(qoff[t]), int(qoff[t + 1])
            ck.build_tree(p_pts[a:b], lay)
            if qb > qa:
                ck.knn(p_pts[a:b], lay, queries[qa:qb], k,
                       out=(p_ids[qa:qb], p_d2[qa:qb], p_r[qa:qb]))

    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.forest_steps):
        plain()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / args.forest_steps

    # batched: forest build + one kNN launch, eager and as a graph
    f = ck.Forest(off, lay)
    b_pts = torch.empty_like(orig)
    b_ids = torch.empty_like(p_ids)
    b_d2 = torch.empty_like(p_d2)
    b_r = torch.empty_like(p_r)

    def batched(stream=None):
        b_pts.copy_(orig)
        f.build(b_pts, stream=stream)
        f.knn(b_pts, queries, k, seg_off, seg_tree, stream=stream, out=(b_ids, b_d2, b_r))

    batched()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.forest_steps):
        batched()
    torch.cuda.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3 / args.forest_steps
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        batched(torch.cuda.current_stream())
    graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.forest_steps):
        graph.replay()
    torch.cuda.synchronize()
    graph_ms = (time.perf_counter() - t0) * 1e3 / args.forest_steps
    same = bool(torch.equal(b_pts, p_pts) and torch.equal(b_ids, p_ids)
                and torch.equal(b_d2, p_d2) and torch.equal(b_r, p_r))
    print("[bench] forest build+knn: plain %.1f ms, batched eager %.1f ms, graph %.1f ms, "
          "identical: %s" % (plain_ms, eager_ms, graph_ms, same), file=sys.stderr, flush=True)
    out = {"forest_trees": T, "forest_points": N, "forest_queries": NQ, "forest_k": k,
           "forest_plain_ms": round(plain_ms, 3), "forest_batched_eager_ms": round(eager_ms, 3),
           "forest_batched_graph_ms": round(graph_ms, 3),
           "batched_vs_plain_speedup": round(plain_ms / graph_ms, 3),
           "batched_equals_plain": same}
    f.close()
    del p_pts, b_pts, orig, queries, p_ids, p_d2, p_r, b_ids, b_d2, b_r
    torch.cuda.empty_cache()
    return out


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
    ap.add_argument("--tl-steps", type=int, default=2,
                    help="timed builds of the bitonic and in-place builders")
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--forest-trees", type=int, default=4096)
    ap.add_argument("--forest-steps", type=int, default=2)
    ap.add_argument("--no-forest", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="for hardware-counter runs: one 1 GiB device copy (the counters' "
                         "calibration), one build, one kNN and one FCP launch, no JSON line")
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

    if args.probe:
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        tree.copy_(pts)
        ck.buildTree(tree, lay)
        ids = torch.empty((nq, args.k), dtype=torch.int32, device=dev)
        d2 = torch.empty((nq, args.k), dtype=torch.float32, device=dev)
        ck.knn(tree, lay, queries, args.k, out=(ids, d2, None), flags=ck.QUERY_SORT)
        ck.fcp(tree, lay, queries, flags=ck.QUERY_SORT)
        torch.cuda.synchronize()
        print("[bench] probe done", file=sys.stderr, flush=True)
        return

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

    # the bitonic (n-int temporary) and in-place (no temporary) builders
    other_builds = {}
    default_tree = tree.clone()
    for name, builder in (("bitonic", ck.BUILDER_BITONIC), ("inplace", ck.BUILDER_INPLACE)):
        times = []
        for _ in range(max(1, args.tl_steps)):
            tree.copy_(pts)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            ck.build_tree(tree, lay, builder)
            e1.record(stream)
            torch.cuda.synchronize()
            times.append(e0.elapsed_time(e1))
        other_builds[name] = (min(times), bool(torch.equal(tree, default_tree)))
        print("[bench] build 10M %s: best %.3f ms, same tree as default: %s"
              % (name, other_builds[name][0], other_builds[name][1]), file=sys.stderr, flush=True)
    tree.copy_(default_tree)
    del default_tree

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

    # ---------------- batched forest vs plain per-tree launches ----------------
    forest = None if args.no_forest else forest_bench(args, ck, datagen, torch, np)

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
            "build_ms_10M_bitonic": round(other_builds["bitonic"][0], 3),
            "build_ms_10M_inplace": round(other_builds["inplace"][0], 3),
            "bitonic_tree_equals_default": other_builds["bitonic"][1],
            "knn_k8_queries_per_s": round(qps, 1),
            "knn_k8_queries_per_s_unordered": round(nq / (plain_ms / 1e3), 1),
            "fcp_queries_per_s": round(nq / (fcp_ms / 1e3), 1),
            "batched_vs_plain_speedup": forest["batched_vs_plain_speedup"] if forest else None,
            "p99_ms": None,
            "visits_per_query": round(visits / nq, 2),
            "forest": forest,
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
