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


def serve_bench(args, ck, datagen, torch, np):
    """Serving loop over resident trees: 512 float3 trees, sizes log-uniform
    in [10K, 2M] (about 190M points), each a 64-centre Gaussian mixture in
    its own random box.  Each request names a tree (uniform) and carries
    1..100K (log-uniform) queries -- data points of that tree plus a small
    Gaussian jitter -- for FCP or kNN with k in {1, 4, 16}.  Saturation
    throughput is measured first, batched and with one launch sequence per
    request (all requests of a burst submitted at once); requests then
    arrive as a Poisson process at a fixed rate (--serve-rate), and the same
    arrival trace is replayed against the batching server and against one
    launch sequence per request.  Latency = submit -> results in the
    caller's host buffers."""
    dev = torch.device("cuda:0")
    T = args.serve_trees
    r = (datagen.splitmix64(71, T) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    sizes = np.exp(np.log(10_000) + r * (np.log(2_000_000) - np.log(10_000))).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N = int(off[-1])
    g = torch.Generator(device=dev)
    g.manual_seed(73)
    lo = torch.rand((T, 3), generator=g, device=dev) * 100.0
    ext = torch.rand((T, 3), generator=g, device=dev) * 10.0 + 0.1
    centres = torch.rand((T, 64, 3), generator=g, device=dev) * ext[:, None, :] + lo[:, None, :]
    tid = torch.repeat_interleave(torch.arange(T, device=dev), torch.from_numpy(sizes).to(dev))
    cid = torch.randint(0, 64, (N,), generator=g, device=dev)
    pts = (centres[tid, cid] + torch.randn((N, 3), generator=g, device=dev)
           * (0.01 * ext[tid])).contiguous()
    del tid, cid
    # the request trace: tree, size, kind; queries = jittered data points
    R = args.serve_requests
    rr = datagen.splitmix64(79, 4 * R)
    u = (rr >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    expo = -np.log(1.0 - u[0::4])
    trees = (rr[1::4] % np.uint64(T)).astype(np.int64)
    nqs = np.exp(u[2::4] * np.log(100_000)).astype(np.int64).clip(1, 100_000)
    kinds = [0, 1, 4, 16]
    ks = np.array([kinds[int(x)] for x in (rr[3::4] % np.uint64(4))])
    qtree = torch.from_numpy(np.repeat(trees, nqs)).to(dev)
    tsz = torch.from_numpy(sizes).to(dev)[qtree]
    pick = torch.minimum((torch.rand(int(nqs.sum()), generator=g, device=dev) * tsz).long(),
                         tsz - 1)
    del tsz
    qidx = torch.from_numpy(off[:-1]).to(dev)[qtree] + pick
    qall = (pts[qidx] + torch.randn((len(qidx), 3), generator=g, device=dev)
            * (0.001 * ext[qtree])).cpu().numpy()
    del qtree, pick, qidx
    qoff = np.concatenate([[0], np.cumsum(nqs)])
    reqs = [(int(trees[i]), qall[qoff[i]:qoff[i + 1]], int(ks[i])) for i in range(R)]
    lay = ck.float_points(3)
    f = ck.Forest(off, lay)
    f.build(pts)
    torch.cuda.synchronize()

    def replay(max_per_batch, arrivals, count):
        srv = ck.Server(f, pts, budget_bytes=args.serve_budget_mb << 20,
                        max_requests_per_batch=max_per_batch)
        w = srv.submit(0, reqs[0][1][:16], 4)  # warm-up
        srv.poll(wait_for=1)
        srv.result(w)
        # result buffers allocated and touched before the clock starts, as a
        # client reusing its buffers would have them
        outs = []
        for i in range(count):
            n = len(reqs[i][1]) * max(1, reqs[i][2])
            outs.append((np.zeros(n, np.int32), np.zeros(n, np.float32)))
            outs[-1][0].fill(1)
            outs[-1][1].fill(1.0)
        tickets = []
        t0 = time.perf_counter()
        for i in range(count):
            tree, q, k = reqs[i]
            if arrivals is not None:
                target = t0 + arrivals[i]
                while time.perf_counter() < target:
                    pass
            tickets.append(srv.submit(tree, q, k, out=outs[i]))
        lat = {}
        while len(lat) < count:
            for c in srv.poll(wait_for=1):
                assert c.status == 0
                lat[c.ticket] = c.total_ms
        wall = time.perf_counter() - t0
        res = [srv.result(t) for t in tickets]
        st = srv.stats()
        srv.close()
        return np.array([lat[t] for t in tickets]), wall, res, int(st.batches)

    # saturation: every request of a burst queued at once, batched and one
    # launch sequence per request
    nb = min(R, 1000)
    _, burst_wall, _, _ = replay(0, None, nb)
    cap_qps = float(nqs[:nb].sum()) / burst_wall
    npb = min(R, 300)
    _, pr_burst_wall, _, _ = replay(1, None, npb)
    pr_cap_qps = float(nqs[:npb].sum()) / pr_burst_wall
    # the Poisson trace at a fixed request rate, against both
    rate = float(args.serve_rate)
    arrivals = np.cumsum(expo / rate)
    b_ms, b_wall, b_res, b_batches = replay(0, arrivals, R)
    npr = int(min(R, max(50, rate * 1.0)))
    p_ms, p_wall, p_res, p_batches = replay(1, arrivals, npr)
    same = all(np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])
               for a, c in zip(b_res[:npr], p_res))
    out = {"serve_trees": T, "serve_points": N, "serve_requests": R,
           "serve_queries": int(nqs.sum()),
           "serve_saturation_queries_per_s": round(cap_qps, 1),
           "per_request_saturation_queries_per_s": round(pr_cap_qps, 1),
           "serve_vs_per_request_saturation": round(cap_qps / pr_cap_qps, 2),
           "serve_rate_req_per_s": round(rate, 1),
           "serve_offered_queries_per_s": round(rate * float(nqs.mean()), 1),
           "serve_p50_ms": round(float(np.percentile(b_ms, 50)), 3),
           "serve_p99_ms": round(float(np.percentile(b_ms, 99)), 3),
           "serve_queries_per_s": round(float(nqs.sum()) / b_wall, 1),
           "serve_batches": b_batches,
           "per_request_requests": npr,
           "per_request_p50_ms": round(float(np.percentile(p_ms, 50)), 3),
           "per_request_p99_ms": round(float(np.percentile(p_ms, 99)), 3),
           "per_request_queries_per_s": round(float(nqs[:npr].sum()) / p_wall, 1),
           "serve_equals_per_request": bool(same)}
    print("[bench] serve: saturation %.1fM q/s batched vs %.1fM q/s per request; at %.0f req/s "
          "batched p50 %.2f ms p99 %.2f ms (%d batches); per-request p50 %.2f ms p99 %.2f ms; "
          "identical: %s"
          % (cap_qps / 1e6, pr_cap_qps / 1e6, rate, out["serve_p50_ms"], out["serve_p99_ms"],
             b_batches, out["per_request_p50_ms"], out["per_request_p99_ms"], same),
          file=sys.stderr, flush=True)
    f.close()
    del pts
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
    ap.add_argument("--serve-trees", type=int, default=512)
    ap.add_argument("--serve-requests", type=int, default=6000)
    ap.add_argument("--serve-budget-mb", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=2000.0,
                    help="Poisson request rate (requests/s) of the serving trace")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="for hardware-counter runs: one 1 GiB device copy (the counters' "
                         "calibration), one build, one kNN and one FCP launch, no JSON line")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path; run with --gpus 1",
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
        # the queries in their Morton launch order, so that the kNN and FCP
        # dispatches are exactly the ones the roofline figure times
        qs = queries[ck.query_order(queries).long()].contiguous()
        ck.knn(tree, lay, qs, args.k, out=(ids, d2, None), flags=ck.QUERY_PRESORTED)
        ck.fcp(tree, lay, qs, flags=ck.QUERY_PRESORTED)
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
    # the kNN kernel alone: the same queries already in the Morton order the
    # step launches them in (and launched with its shape), event-timed on
    # the launch stream
    qs = queries[ck.query_order(queries).long()].contiguous()
    for _ in range(args.warmup):
        ck.knn(tree, lay, qs, k, out=out, flags=ck.QUERY_PRESORTED)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        ck.knn(tree, lay, qs, k, out=out, flags=ck.QUERY_PRESORTED)
    e1.record(stream)
    torch.cuda.synchronize()
    kernel_ms = e0.elapsed_time(e1) / args.steps
    # the same launch with one query per lane (the kernel the default
    # two-per-lane one replaces), for comparison
    for _ in range(args.warmup):
        ck.knn(tree, lay, qs, k, out=out, flags=ck.QUERY_ONE_PER_LANE | ck.QUERY_PRESORTED)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        ck.knn(tree, lay, qs, k, out=out, flags=ck.QUERY_ONE_PER_LANE | ck.QUERY_PRESORTED)
    e1.record(stream)
    torch.cuda.synchronize()
    one_lane_ms = e0.elapsed_time(e1) / args.steps
    del qs
    knn_gbs = knn_bytes / (kernel_ms / 1e3) / 1e9
    print("[bench] knn k=%d: %.3f ms/step (events %.3f ms, kernel alone %.3f ms), visits "
          "%.1f/query" % (k, ms_per_step, ev_ms, kernel_ms, visits / nq),
          file=sys.stderr, flush=True)

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
    serve = None if args.no_serve else serve_bench(args, ck, datagen, torch, np)

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

    # HBM bytes per kNN launch from the committed hardware-counter summary of
    # this same workload (FETCH_SIZE + WRITE_SIZE, each scaled by what the
    # counter reads for a 1 GiB device copy)
    traffic = None
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_knn_10M_float3_k8.json")) as fh:
            traffic = json.load(fh)["traffic_bytes_per_launch"]
    except (OSError, ValueError, KeyError):
        pass

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
            "knn_k8_kernel_ms_one_query_per_lane": round(one_lane_ms, 4),
            "fcp_queries_per_s": round(nq / (fcp_ms / 1e3), 1),
            "batched_vs_plain_speedup": forest["batched_vs_plain_speedup"] if forest else None,
            "p99_ms": serve["serve_p99_ms"] if serve else None,
            "visits_per_query": round(visits / nq, 2),
            "forest": forest,
            "serve": serve,
        },
        "roofline": {"bound": "hbm", "achieved": round(knn_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(knn_gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "kernel": "k_query2<3, FixedCandidateList<8>> (stack-based, two queries "
                               "per lane) on Morton-ordered queries",
                     "algorithmic_bytes_per_launch": knn_bytes,
                     "avg_launch_ms": round(kernel_ms, 4),
                     "step_ms_with_ordering": round(ev_ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
