"""Benchmark of the build-and-query path on one MI355X.

Workload (the 10M-point float3 case of the builder/query scaling config):
  * tree: 10M float3 points, uniform in [0,1)^3 (SplitMix64 seed 1),
    left-balanced round-robin tree built by cukd_build_tree;
  * step: one kNN (k=8, FixedCandidateList, stack-based traversal) batch
    of 1M uniform float3 queries (seed 2) over the resident tree.

Prints one JSON line.  value = queries/s of the kNN step; the build time,
FCP rate and the other headline figures ride along under "metrics", the
small plain-method case, the rest of the builder/query scaling config
(stack-free queries, 1M and GMM trees) and the spatial k-d tree case
under "metrics.c1", "metrics.c2" and "metrics.spatial", the forest and
serving forms under
"metrics.forest" and "metrics.serve".  Inputs are resident in HBM before
any timed region starts.  "parity" records the checks this run made
against the oracle (test infrastructure, never the measured path): the
10M tree byte for byte, and the headline launch's results on a 100K-query
sample.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
"""
import argparse
import ctypes
import hashlib
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# source files that define the headline kNN kernel; the committed counter
# summary records their hash, and its traffic figure is used only while
# the hash still matches
KNN_KERNEL_SOURCES = ("cudakdtree_amd/csrc/query_impl.h", "cudakdtree_amd/csrc/query_inst.hip",
                      "cudakdtree_amd/csrc/cukd_common.h", "include/cukd/traverse.h",
                      "include/cukd/knn.h", "include/cukd/fcp.h", "cudakdtree_amd/build.py")


def kernel_source_sha():
    h = hashlib.sha256()
    for rel in KNN_KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(rel.encode() + b"\0" + f.read())
    return h.hexdigest()[:16]


def host_cores():
    """(threads used for the CPU baseline, nproc, CPUs this process may run
    on).  A one-GPU job's share of the GPU box's host is 16 cores; the
    baseline uses that share (or fewer if the affinity mask is smaller)."""
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = nproc
    return max(1, min(16, allowed)), nproc, allowed


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

    if args.probe_forest:
        f = ck.Forest(off, lay)
        b_pts = torch.empty_like(orig)
        b_ids = torch.empty((NQ, k), dtype=torch.int32, device=dev)
        b_d2 = torch.empty((NQ, k), dtype=torch.float32, device=dev)
        b_r = torch.empty(NQ, dtype=torch.float32, device=dev)

        def step(stream=None):
            b_pts.copy_(orig)
            f.build(b_pts, stream=stream)
            f.knn(b_pts, queries, k, seg_off, seg_tree, stream=stream, out=(b_ids, b_d2, b_r),
                  flags=ck.QUERY_AUTOTUNE)

        step()
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            step(torch.cuda.current_stream())
        for i in range(3):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            graph.replay()
            e1.record()
            torch.cuda.synchronize()
            print("[bench] forest graph replay %d: %.2f ms" % (i, e0.elapsed_time(e1)),
                  file=sys.stderr, flush=True)
        f.close()
        return None

    # plain: one cukd_build_tree + one cukd_knn per tree, the loop in C++
    # (cudakdtree_amd/cli/plain_forest.cpp), autotuned launch shapes
    p_pts = torch.empty_like(orig)
    p_ids = torch.empty((NQ, k), dtype=torch.int32, device=dev)
    p_d2 = torch.empty((NQ, k), dtype=torch.float32, device=dev)
    p_r = torch.empty(NQ, dtype=torch.float32, device=dev)
    plain_lib = ctypes.CDLL(os.path.join(ROOT, "cudakdtree_amd", "cli", "libcukd_plain.so"))
    ms = ctypes.c_double(0.0)
    h_off = np.ascontiguousarray(off, dtype=np.int64)
    h_qoff = np.ascontiguousarray(qoff, dtype=np.int64)
    cs = torch.cuda.current_stream()
    rc = plain_lib.cukd_plain_forest_run(
        ctypes.c_void_p(orig.data_ptr()), ctypes.c_void_p(p_pts.data_ptr()),
        h_off.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(T),
        ctypes.c_void_p(queries.data_ptr()), h_qoff.ctypes.data_as(ctypes.c_void_p),
        ctypes.c_int32(k), ctypes.c_void_p(p_ids.data_ptr()), ctypes.c_void_p(p_d2.data_ptr()),
        ctypes.c_void_p(p_r.data_ptr()), ctypes.c_int32(ck.QUERY_AUTOTUNE),
        ctypes.c_void_p(cs.cuda_stream), ctypes.c_int32(args.forest_steps), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError("plain forest run failed: %s" % ck.lib().cukd_last_error().decode())
    plain_ms = ms.value

    # batched: forest build + one kNN launch, eager and as a graph; the kNN
    # launch shape is autotuned on the first (eager) call, before capture
    f = ck.Forest(off, lay)
    b_pts = torch.empty_like(orig)
    b_ids = torch.empty_like(p_ids)
    b_d2 = torch.empty_like(p_d2)
    b_r = torch.empty_like(p_r)

    def batched(stream=None):
        b_pts.copy_(orig)
        f.build(b_pts, stream=stream)
        f.knn(b_pts, queries, k, seg_off, seg_tree, stream=stream, out=(b_ids, b_d2, b_r),
              flags=ck.QUERY_AUTOTUNE)

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
           "forest_plain": "one cukd_build_tree + one cukd_knn per tree, looped in C++ "
                           "(cudakdtree_amd/cli/plain_forest.cpp); both legs autotuned",
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
    nb = min(R, 3000)
    _, burst_wall, _, _ = replay(0, None, nb)
    cap_qps = float(nqs[:nb].sum()) / burst_wall
    npb = min(R, 1000)
    _, pr_burst_wall, _, _ = replay(1, None, npb)
    pr_cap_qps = float(nqs[:npb].sum()) / pr_burst_wall
    # the Poisson trace at a fixed request rate, against both
    rate = float(args.serve_rate)
    arrivals = np.cumsum(expo / rate)
    b_ms, b_wall, b_res, b_batches = replay(0, arrivals, R)
    # the same trace, every request, one launch sequence each: both tails
    # are taken over the identical requests
    npr = R
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


def c2_bench(args, ck, datagen, torch, np, tree, lay, queries):
    """The rest of the builder/query scaling config: stack-free FCP and
    kNN-8 over the 10M uniform tree, and the 1M uniform and 10M GMM (64
    centres, sigma 0.01) trees, each with its build ms (median of 5) and
    stack-based FCP / kNN-8 rates; always the same 1M uniform queries
    (seed 2), Morton-ordered launches (ordering time included)."""
    dev = queries.device
    nq = queries.shape[0]
    k = 8
    ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    d2 = torch.empty((nq, k), dtype=torch.float32, device=dev)
    rr = torch.empty(nq, dtype=torch.float32, device=dev)
    fi = torch.empty(nq, dtype=torch.int32, device=dev)
    fd = torch.empty(nq, dtype=torch.float32, device=dev)
    def rate(fn, steps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return round(nq / ((time.perf_counter() - t0) / steps), 1)

    def rates(t, traversal, steps):
        knn = lambda: ck.knn(t, lay, queries, k, traversal=traversal, out=(ids, d2, rr),
                             flags=ck.QUERY_SORT)
        counter = torch.zeros(1, dtype=torch.int64, device=dev)
        ck.set_traversal_stats(counter)
        knn()
        torch.cuda.synchronize()
        ck.set_traversal_stats(None)
        return {
            "knn_k8_queries_per_s": rate(knn, steps),
            "knn_k8_visits_per_query": round(int(counter.item()) / nq, 2),
            "fcp_queries_per_s": rate(lambda: ck.fcp(t, lay, queries, traversal=traversal,
                                                     out=(fi, fd), flags=ck.QUERY_SORT), steps)}

    out = {"workload": "1M uniform float3 queries (seed 2) over: the 10M uniform tree with the "
                       "stack-free traversal; 1M uniform (seed 1) and 10M GMM (64 centres, "
                       "sigma 0.01, seed 1) trees with the stack-based one",
           "stack_free_10M_uniform": rates(tree, ck.STACK_FREE, max(3, args.steps // 2))}
    stream = torch.cuda.current_stream()
    for name, pts_h in (("uniform_1M", datagen.uniform(1_000_000, 3, 1)),
                        ("gmm_10M", datagen.gmm(10_000_000, 3, 1))):
        p = torch.from_numpy(pts_h).to(dev)
        t = torch.empty_like(p)
        ms = []
        for i in range(6):
            t.copy_(p)
            torch.cuda.synchronize()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            ck.buildTree(t, lay)
            e1.record(stream)
            torch.cuda.synchronize()
            if i > 0:
                ms.append(e0.elapsed_time(e1))
        # uniform queries over the clustered tree are mostly far from every
        # cluster and visit many nodes each: fewer timed repetitions
        reps = 2 if name == "gmm_10M" else max(3, args.steps // 2)
        out[name] = {"build_ms": round(_pct(ms, 50), 3), **rates(t, ck.STACK_BASED, reps)}
        del p, t
    print("[bench] c2: %s" % json.dumps({kk: v for kk, v in out.items() if kk != "workload"}),
          file=sys.stderr, flush=True)
    return out


def lists_bench(args, ck, torch, tree, lay, queries):
    """Candidate lists beyond the register-resident FixedCandidateList<k>
    (k = 1/4/8/16): stack-based kNN of the 1M queries over the 10M tree,
    Morton-ordered once outside the timed region, for the fixed list at
    k = 16 (registers), 20 and 64 and the heap list at k = 16, 20, 50 and 64
    (run-time k, entries in LDS).  Full-size consistency: the sorted
    heap-64 keys equal the fixed-64 list, their first 20 the fixed-20 list,
    whose first 16 equal the register list, and every heap row is a
    max-heap."""
    dev = queries.device
    nq = queries.shape[0]
    qs = queries[ck.query_order(queries).long()].contiguous()
    stream = torch.cuda.current_stream()
    res, keys, heaps = {}, {}, {}
    for lk, kk in ((ck.LIST_FIXED, 16), (ck.LIST_FIXED, 20), (ck.LIST_FIXED, 64),
                   (ck.LIST_HEAP, 16), (ck.LIST_HEAP, 20), (ck.LIST_HEAP, 50),
                   (ck.LIST_HEAP, 64)):
        name = "%s_k%d" % ("fixed" if lk == ck.LIST_FIXED else "heap", kk)
        ids = torch.empty((nq, kk), dtype=torch.int32, device=dev)
        d2 = torch.empty((nq, kk), dtype=torch.float32, device=dev)
        rr = torch.empty(nq, dtype=torch.float32, device=dev)
        fn = lambda: ck.knn(tree, lay, qs, kk, list_kind=lk, out=(ids, d2, rr),
                            flags=ck.QUERY_PRESORTED)
        fn()
        torch.cuda.synchronize()
        reps = 3
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": round(ms, 3), "queries_per_s": round(nq / (ms / 1e3), 1)}
        if name in ("fixed_k16", "fixed_k20", "fixed_k64", "heap_k64"):
            key = (d2.view(torch.int32).to(torch.int64) << 32) | (ids.to(torch.int64) & 0xFFFFFFFF)
            (heaps if lk == ck.LIST_HEAP else keys)[kk] = key
        del ids, d2, rr
    h = heaps[64]
    par = torch.arange(1, 64, device=dev).sub_(1).div_(2, rounding_mode="floor")
    is_heap = bool((h[:, par] >= h[:, 1:]).all().item())
    hs = torch.sort(h, dim=1).values
    consistent = (is_heap and torch.equal(hs, keys[64]) and torch.equal(hs[:, :20], keys[20])
                  and torch.equal(keys[20][:, :16], keys[16]))
    out = {"workload": "stack-based kNN of 1M uniform float3 queries (seed 2, Morton-ordered "
                       "once) over the 10M-point tree; fixed k=16 is the register list",
           **res, "lists_consistent": consistent}
    print("[bench] lists: %s" % json.dumps(out), file=sys.stderr, flush=True)
    del keys, heaps, h, hs, qs
    return out


def multiround_bench(args, ck, torch, pts, lay, queries, tree):
    """Multi-round kNN (samples/mpiHugeQuery.cu:75-103): the 10M points
    split into 4 contiguous shards of 2.5M, each built as its own tree (one
    GPU's share of a 4-rank ring, every shard visiting in turn), and 4
    rounds of stack-free kNN-8 of the 1M queries into persistent heap lists
    in HBM (Morton-ordered launches, global ids = shard offset + index),
    then the extraction.  Full-size checks: the extracted dist2 equal the
    single-tree kNN-8 over all 10M points, and every id's point lies at
    its reported dist2."""
    dev = queries.device
    nq, k, P = queries.shape[0], 8, 4
    n = pts.shape[0]
    off = [i * n // P for i in range(P + 1)]
    shards = []
    for r in range(P):
        t = pts[off[r]:off[r + 1]].clone()
        ck.buildTree(t, lay)
        shards.append(t)
    lists = ck.knn_lists(nq, k, device=dev)

    def run():
        for r in range(P):
            ck.knn_accumulate(shards[r], lay, queries, lists, init=(r == 0), id_offset=off[r],
                              flags=ck.QUERY_SORT)
        return ck.knn_extract(lists)

    run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    reps = 3
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        ids, d2, rad = run()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    _, fd2, _ = ck.knn(tree, lay, queries, k, traversal=ck.STACK_FREE, flags=ck.QUERY_SORT)
    allp = torch.cat(shards)
    p = allp[ids.long()]
    dd = p - queries[:, None, :]
    own = (dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1]) + dd[..., 2] * dd[..., 2]
    ok = bool(torch.equal(d2, fd2) and (ids >= 0).all().item() and torch.equal(own, d2))
    out = {"workload": "1M uniform float3 queries (seed 2), kNN k=8, over the 10M uniform points "
                       "as 4 shards of 2.5M: 4 accumulate rounds (stack-free, Morton-ordered) "
                       "+ extract",
           "shards": P, "k": k, "ms_4_rounds_and_extract": round(ms, 3),
           "queries_per_s": round(nq / (ms / 1e3), 1),
           "equals_single_tree_knn": ok}
    print("[bench] multiround: %s" % json.dumps(out), file=sys.stderr, flush=True)
    del shards, lists, allp, p, dd, own
    return out


def c1_bench(args, ck, datagen, torch, np):
    """The plain method at its small size: a left-balanced tree over 100K
    float3 points uniform in [0,1)^3 (seed 1) built by buildTree, then FCP
    (stack-based) for 100K uniform queries (seed 2).  GPU: event-timed
    build (median of 10) and FCP launch; CPU: the oracle's buildTree_host
    (one core) and its host FCP on the allowed cores."""
    dev = torch.device("cuda:0")
    lay = ck.float_points(3)
    n = nq = 100_000
    pts_h = datagen.uniform(n, 3, 1)
    q_h = datagen.uniform(nq, 3, 2)
    src = torch.from_numpy(pts_h).to(dev)
    q = torch.from_numpy(q_h).to(dev)
    t = torch.empty_like(src)
    ids = torch.empty(nq, dtype=torch.int32, device=dev)
    d2 = torch.empty(nq, dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream()
    b_ms, f_ms = [], []
    for i in range(12):
        t.copy_(src)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e2 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        ck.buildTree(t, lay)
        e1.record(s)
        ck.fcp(t, lay, q, out=(ids, d2))
        e2.record(s)
        torch.cuda.synchronize()
        if i >= 2:
            b_ms.append(e0.elapsed_time(e1))
            f_ms.append(e1.elapsed_time(e2))
    out = {"workload": "100K float3 U[0,1)^3 buildTree + 100K FCP (stack-based)",
           "gpu_build_ms": round(_pct(b_ms, 50), 4), "gpu_fcp_ms": round(_pct(f_ms, 50), 4),
           "gpu_fcp_queries_per_s": round(nq / (_pct(f_ms, 50) / 1e3), 1)}
    if not args.no_cpu:
        import oracle as oc
        threads = host_cores()[0]
        t0 = time.perf_counter()
        ref, _ = oc.build_host(pts_h, lay)
        cb = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        rid, rd2 = oc.fcp(oc.STACK_BASED, ref, lay, q_h, threads=threads)
        cf = (time.perf_counter() - t0) * 1e3
        out.update({"cpu_build_ms_single_core": round(cb, 2), "cpu_fcp_ms": round(cf, 2),
                    "cpu_fcp_cores": threads,
                    "_tree_ok": bool(np.array_equal(ref, t.cpu().numpy())),
                    "_fcp_ok": bool(np.array_equal(rid, ids.cpu().numpy()) and np.array_equal(
                        rd2.view(np.uint32), d2.cpu().numpy().view(np.uint32)))})
    print("[bench] c1 100K: build %.3f ms, fcp %.3f ms" % (out["gpu_build_ms"], out["gpu_fcp_ms"]),
          file=sys.stderr, flush=True)
    return out


def spatial_bench(args, ck, datagen, torch, np):
    """The spatial k-d tree case: 4M float4 points uniform in [-1,1)^4 with
    1% exact duplicates; spatial build (the reference's host-synchronised
    round loop), then kNN k=16 with cutoff 0.05 (stack-based,
    FixedCandidateList<16>) for 2M queries, half data points plus
    N(0, 0.001) jitter and half uniform; and a 2M int3 set uniform in
    [0, 2^20) (1% duplicates), build only (the reference has no int
    distance math).  CPU: the oracle's spatial build (one core) and its
    kNN on a 50K-query sample (allowed cores)."""
    dev = torch.device("cuda:0")
    n, D, nq = 4_000_000, 4, 2_000_000
    pts_h = datagen.uniform(n, D, 11, -1.0, 1.0)
    dup = (datagen.splitmix64(12, n // 100) % np.uint64(n)).astype(np.int64)
    pts_h[n - n // 100:] = pts_h[dup]
    half = nq // 2
    pick = (datagen.splitmix64(13, half) % np.uint64(n)).astype(np.int64)
    jit = (datagen.gaussian(half * D, 14).reshape(half, D) * 0.001).astype(np.float32)
    q_h = np.concatenate([pts_h[pick] + jit,
                          datagen.uniform(nq - half, D, 15, -1.0, 1.0)]).astype(np.float32)
    lay = ck.float_points(D)
    pts = torch.from_numpy(pts_h).to(dev)
    q = torch.from_numpy(q_h).to(dev)
    torch.cuda.synchronize()
    bms = []
    tree = None
    for i in range(4):
        if tree is not None:
            ck.free(tree)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tree = ck.buildSpatialTree(pts, lay)  # synchronises, as the reference's build does
        bms.append((time.perf_counter() - t0) * 1e3)
    k, r = 16, 0.05
    ids = torch.empty((nq, k), dtype=torch.int32, device=dev)
    d2 = torch.empty((nq, k), dtype=torch.float32, device=dev)
    ret = torch.empty(nq, dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream()
    qms = []
    for i in range(6):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        lib = ck.lib()
        rc = lib.cukd_spatial_knn(ck.STACK_BASED, ck.LIST_FIXED, k, ctypes.byref(tree.raw),
                                  ctypes.byref(lay), ctypes.c_void_p(q.data_ptr()), nq, r,
                                  ctypes.c_void_p(ids.data_ptr()), ctypes.c_void_p(d2.data_ptr()),
                                  ctypes.c_void_p(ret.data_ptr()), ctypes.c_void_p(s.cuda_stream))
        assert rc == 0, lib.cukd_last_error()
        e1.record(s)
        torch.cuda.synchronize()
        if i >= 1:
            qms.append(e0.elapsed_time(e1))
    # the reference's spatial sample keeps a HeapCandidateList<16>
    # (samples/knn-float3-spatialkdtree.cu:26): the heap list (LDS columns),
    # whose sorted keys must equal the fixed list's
    hids = torch.empty_like(ids)
    hd2 = torch.empty_like(d2)
    hms = []
    for i in range(6):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        rc = lib.cukd_spatial_knn(ck.STACK_BASED, ck.LIST_HEAP, k, ctypes.byref(tree.raw),
                                  ctypes.byref(lay), ctypes.c_void_p(q.data_ptr()), nq, r,
                                  ctypes.c_void_p(hids.data_ptr()), ctypes.c_void_p(hd2.data_ptr()),
                                  ctypes.c_void_p(ret.data_ptr()), ctypes.c_void_p(s.cuda_stream))
        assert rc == 0, lib.cukd_last_error()
        e1.record(s)
        torch.cuda.synchronize()
        if i >= 1:
            hms.append(e0.elapsed_time(e1))
    key = lambda a, b: (b.view(torch.int32).to(torch.int64) << 32) | (a.to(torch.int64) & 0xFFFFFFFF)
    heap_ok = torch.equal(torch.sort(key(hids, hd2), dim=1).values, key(ids, d2))
    del hids, hd2
    # int3 set, build only
    ni = 2_000_000
    ci = (datagen.splitmix64(17, ni * 3) >> np.uint64(44)).astype(np.int32).reshape(ni, 3)
    src = (datagen.splitmix64(18, ni // 100) % np.uint64(ni)).astype(np.int64)
    ci[ni - ni // 100:] = ci[src]
    ilay = ck.int_points(3)
    dci = torch.from_numpy(ci).to(dev)
    ims = []
    itree = None
    for i in range(4):
        if itree is not None:
            ck.free(itree)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        itree = ck.buildSpatialTree(dci, ilay)
        ims.append((time.perf_counter() - t0) * 1e3)
    out = {"workload": "spatial k-d tree: 4M float4 U[-1,1)^4 (1% duplicates) build + kNN k=16 "
                       "r=0.05 (stack-based, fixed and heap lists) of 2M queries; 2M int3 build",
           "gpu_build_ms_4M_float4": round(min(bms[1:]), 3),
           "gpu_knn16_ms_2M": round(_pct(qms, 50), 3),
           "gpu_knn16_queries_per_s": round(nq / (_pct(qms, 50) / 1e3), 1),
           "gpu_knn16_heap_ms_2M": round(_pct(hms, 50), 3),
           "knn16_heap_sorted_equals_fixed": heap_ok,
           "gpu_build_ms_2M_int3": round(min(ims[1:]), 3), "nodes_4M": tree.num_nodes}
    if not args.no_cpu:
        import oracle as oc
        threads = host_cores()[0]
        t0 = time.perf_counter()
        rn, rp, rb = oc.spatial_build(pts_h, lay)
        cb = (time.perf_counter() - t0) * 1e3
        build_ok = bool(np.array_equal(tree.nodes_numpy().view(np.uint8), rn.view(np.uint8))
                        and np.array_equal(tree.prims_numpy(), rp))
        m = 50_000
        sel = np.concatenate([np.arange(m // 2), half + np.arange(m - m // 2)])
        t0 = time.perf_counter()
        cid, cd2, cret = oc.spatial_knn(oc.STACK_BASED, oc.LIST_FIXED, k, rn, rp, rb, pts_h, lay,
                                        q_h[sel], r, threads=threads)
        cq = time.perf_counter() - t0
        gi, gd = ids.cpu().numpy()[sel], d2.cpu().numpy()[sel]
        knn_ok = bool(np.array_equal(cid, gi) and np.array_equal(cd2.view(np.uint32),
                                                                 gd.view(np.uint32)))
        t0 = time.perf_counter()
        irn, irp, _ = oc.spatial_build(ci, ilay)
        cib = (time.perf_counter() - t0) * 1e3
        int_ok = bool(np.array_equal(itree.nodes_numpy().view(np.uint8), irn.view(np.uint8))
                      and np.array_equal(itree.prims_numpy(), irp))
        out.update({"cpu_build_ms_4M_float4_single_core": round(cb, 1),
                    "cpu_knn16_queries_per_s": round(m / cq, 1), "cpu_knn16_cores": threads,
                    "cpu_knn16_sample": "%d of the 2M queries (half from each half)" % m,
                    "cpu_build_ms_2M_int3_single_core": round(cib, 1),
                    "_spatial_build": build_ok, "_spatial_knn_sample": knn_ok,
                    "_spatial_int3_build": int_ok})
    ck.free(tree)
    ck.free(itree)
    print("[bench] spatial: 4M float4 build %.1f ms, knn16 %.3f ms for 2M queries, 2M int3 "
          "build %.1f ms" % (out["gpu_build_ms_4M_float4"], out["gpu_knn16_ms_2M"],
                            out["gpu_build_ms_2M_int3"]), file=sys.stderr, flush=True)
    del pts, q, ids, d2, ret, dci
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
    ap.add_argument("--serve-requests", type=int, default=12000)
    ap.add_argument("--serve-budget-mb", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=2000.0,
                    help="Poisson request rate (requests/s) of the serving trace")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-lists", action="store_true")
    ap.add_argument("--no-multiround", action="store_true")
    ap.add_argument("--no-spatial", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="for hardware-counter runs: one 1 GiB device copy (the counters' "
                         "calibration), one build, one kNN and one FCP launch, no JSON line")
    ap.add_argument("--probe-build", action="store_true",
                    help="for kernel-trace runs: five builds of the 10M tree, no JSON line")
    ap.add_argument("--probe-build-explicit", action="store_true",
                    help="for kernel-trace runs: five builds of the 10M explicit-dim "
                         "PointPlusPayload tree, no JSON line")
    ap.add_argument("--probe-forest", action="store_true",
                    help="for kernel-trace runs: the forest step eager once, then three "
                         "replays of its graph, no JSON line")
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

    if args.probe_forest:
        forest_bench(args, ck, datagen, torch, np)
        print("[bench] probe-forest done", file=sys.stderr, flush=True)
        return

    if args.probe_build_explicit:
        lay_x = ck.point_plus_payload(3, explicit_dim=True)
        px = torch.zeros((n, 5), dtype=torch.int32, device=dev)
        px[:, :3] = pts.view(torch.int32).reshape(n, 3)
        px[:, 3] = torch.arange(n, dtype=torch.int32, device=dev)
        tx = torch.empty_like(px)
        wbx = ck.bounds_tensor(lay_x)
        for _ in range(5):
            tx.copy_(px)
            ck.buildTree(tx, lay_x, wbx)
        torch.cuda.synchronize()
        print("[bench] probe-build-explicit done", file=sys.stderr, flush=True)
        return

    if args.probe_build:
        for _ in range(5):
            tree.copy_(pts)
            ck.buildTree(tree, lay)
        torch.cuda.synchronize()
        print("[bench] probe-build done", file=sys.stderr, flush=True)
        return

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

    # explicit split dims (widest extent per node): PointPlusPayload
    # {float3 point; int payload; int dim}, 20 bytes, over the same points
    # with payload = input index; compared with the bitonic builder's tree
    lay_x = ck.point_plus_payload(3, explicit_dim=True)
    px = torch.zeros((n, 5), dtype=torch.int32, device=dev)
    px[:, :3] = pts.view(torch.int32).reshape(n, 3)
    px[:, 3] = torch.arange(n, dtype=torch.int32, device=dev)
    tx = torch.empty_like(px)
    wbx = ck.bounds_tensor(lay_x)
    expl_ms = []
    for i in range(args.build_steps + 1):
        tx.copy_(px)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        ck.buildTree(tx, lay_x, wbx)
        e1.record(stream)
        torch.cuda.synchronize()
        if i > 0:
            expl_ms.append(e0.elapsed_time(e1))
    tb = px.clone()
    ck.build_tree(tb, lay_x, ck.BUILDER_BITONIC, ck.bounds_tensor(lay_x))
    torch.cuda.synchronize()
    expl_same = bool(torch.equal(tx, tb))
    expl_med = _pct(expl_ms, 50)
    del px, tx, tb
    print("[bench] build 10M explicit-dim payload: median %.3f ms, same tree as bitonic: %s"
          % (expl_med, expl_same), file=sys.stderr, flush=True)

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
    # the headline launch's results on the CPU sample, for the parity check
    head_ids = ids[: args.cpu_sample].cpu().numpy()
    head_d2 = d2[: args.cpu_sample].cpu().numpy()
    head_ret = ret[: args.cpu_sample].cpu().numpy()

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

    # ---------------- the other configurations' legs ----------------
    c2 = None if args.no_c2 else c2_bench(args, ck, datagen, torch, np, tree, lay, queries)
    lists = None if args.no_lists else lists_bench(args, ck, torch, tree, lay, queries)
    multiround = (None if args.no_multiround else
                  multiround_bench(args, ck, torch, pts, lay, queries, tree))
    c1 = None if args.no_c1 else c1_bench(args, ck, datagen, torch, np)
    spatial = None if args.no_spatial else spatial_bench(args, ck, datagen, torch, np)
    forest = None if args.no_forest else forest_bench(args, ck, datagen, torch, np)
    serve = None if args.no_serve else serve_bench(args, ck, datagen, torch, np)

    # ---------------- CPU baseline and parity (oracle, host cores) ----------------
    cpu, parity = None, None
    if not args.no_cpu:
        import oracle as oc
        threads, nproc, allowed = host_cores()
        tree_h = tree.cpu().numpy()
        # the 10M host build, once (buildTree_host is single-threaded), and
        # the GPU tree against it byte for byte
        print("[bench] oracle: 10M host build (single core)", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        ref_tree, _ = oc.build_host(pts_h, lay)
        host_build_ms = (time.perf_counter() - t0) * 1e3
        tree_ok = bool(np.array_equal(ref_tree, tree_h))
        del ref_tree
        # the headline kNN on a sample of its queries, all allowed cores
        qs = q_h[: args.cpu_sample]
        t0 = time.perf_counter()
        rid, rd2, rret = oc.knn(oc.STACK_BASED, oc.LIST_FIXED, k, tree_h, lay, qs,
                                threads=threads)
        cpu_s = time.perf_counter() - t0
        knn_ok = bool(np.array_equal(rid, head_ids)
                      and np.array_equal(rd2.view(np.uint32), head_d2.view(np.uint32))
                      and np.array_equal(rret.view(np.uint32), head_ret.view(np.uint32)))
        cpu = {"value": round(len(qs) / cpu_s, 1), "unit": "queries/s", "cores": threads,
               "nproc": nproc, "affinity_cpus": allowed, "kind": "port",
               "sample": "%d of the %d kNN(k=%d) queries over the same 10M-point tree, oracle "
                         "stack-based traversal, OpenMP over queries on %d threads (the "
                         "one-GPU job's share of the host)" % (len(qs), nq, k, threads),
               "build_ms_10M_single_core": round(host_build_ms, 1),
               "build_note": "oracle buildTree_host restatement at 10M, run once, one core"}
        parity = {"tree_10M": tree_ok, "knn_100K": knn_ok}
        if c1:
            parity.update({"c1_tree": c1.pop("_tree_ok"), "c1_fcp": c1.pop("_fcp_ok")})
        if spatial:
            parity.update({k2: spatial.pop("_" + k2) for k2 in
                           ("spatial_build", "spatial_knn_sample", "spatial_int3_build")})
        print("[bench] parity: %s" % parity, file=sys.stderr, flush=True)

    # HBM bytes per kNN launch from the committed hardware-counter summary of
    # this same workload (FETCH_SIZE + WRITE_SIZE, each scaled by what the
    # counter reads for a 1 GiB device copy), used only if it was measured
    # on this kernel's current sources
    traffic, traffic_note = None, "no counter summary"
    try:
        with open(os.path.join(ROOT, "profiles", "pmc_knn_10M_float3_k8.json")) as fh:
            pmc = json.load(fh)
        if pmc.get("kernel_source_sha") == kernel_source_sha():
            traffic = pmc["traffic_bytes_per_launch"]
            traffic_note = "profiles/pmc_knn_10M_float3_k8.json, kernel sources %s" % \
                pmc["kernel_source_sha"]
        else:
            traffic_note = "counter summary is of other kernel sources (%s, now %s)" % (
                pmc.get("kernel_source_sha"), kernel_source_sha())
    except (OSError, ValueError, KeyError):
        pass
    if c1 is not None:
        c1 = {k2: v for k2, v in c1.items() if not k2.startswith("_")}
    if spatial is not None:
        spatial = {k2: v for k2, v in spatial.items() if not k2.startswith("_")}

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
            "build_ms_10M_explicit_payload": round(expl_med, 3),
            "explicit_tree_equals_bitonic": expl_same,
            "knn_k8_queries_per_s": round(qps, 1),
            "knn_k8_queries_per_s_unordered": round(nq / (plain_ms / 1e3), 1),
            "knn_k8_kernel_ms_one_query_per_lane": round(one_lane_ms, 4),
            "fcp_queries_per_s": round(nq / (fcp_ms / 1e3), 1),
            "batched_vs_plain_speedup": forest["batched_vs_plain_speedup"] if forest else None,
            "p99_ms": serve["serve_p99_ms"] if serve else None,
            "visits_per_query": round(visits / nq, 2),
            "c1": c1,
            "c2": c2,
            "lists": lists,
            "multiround": multiround,
            "spatial": spatial,
            "forest": forest,
            "serve": serve,
        },
        "parity": parity,
        "roofline": {"bound": "hbm", "achieved": round(knn_gbs, 2), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(knn_gbs / HBM_PEAK_GBS, 4),
                     "traffic": traffic,
                     "traffic_source": traffic_note,
                     "kernel": "k_query2s<3, FixedCandidateList<8>> (stack-based, two queries "
                               "per lane, pending plane distances in LDS) on Morton-ordered "
                               "queries",
                     "algorithmic_bytes_per_launch": knn_bytes,
                     "avg_launch_ms": round(kernel_ms, 4),
                     "step_ms_with_ordering": round(ev_ms, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
