"""Forests (cukd_forest_*): many independent trees built in one sequence of
launches and queried in one launch.  Every tree equals the oracle's
buildTree_host over its points (builder_thrust.h:514-601), and every query
result equals the single-tree query on that tree (ids in that tree)."""
import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.helpers import distribution, layout_for, make_elements

pytestmark = pytest.mark.gpu


def log_uniform_sizes(count, lo, hi, seed, empty_every=0):
    r = (datagen.splitmix64(seed, count) >> np.uint64(11)).astype(np.float64) / float(1 << 53)
    n = np.exp(np.log(lo) + r * (np.log(hi) - np.log(lo))).astype(np.int64)
    if empty_every:
        n[::empty_every] = 0
    return n


def offsets_of(sizes):
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def forest_points(sizes, D, seed, kind="float", dist="uniform"):
    """Each tree uniform in its own random axis-aligned box."""
    chunks = []
    for t, n in enumerate(sizes):
        c = distribution(dist, int(n), D, seed=seed + t)
        lo = datagen.uniform(1, D, seed * 7 + t)[0] * 10.0
        c = (c * (1.0 + t % 3) + lo).astype(np.float32)
        chunks.append(c)
    return np.concatenate(chunks) if chunks else np.zeros((0, D), np.float32)


@pytest.mark.parametrize("kind", ["float", "payload", "explicit", "photon"])
@pytest.mark.parametrize("dist", ["default", "grid17", "coarse"])
def test_forest_build_matches_oracle_per_tree(kind, dist):
    """dist 'grid17', 'coarse': short runs of equal coordinates, which the
    rank finish repairs level by level (or, where dense, hands to the
    bitonic finish) where a tree's finish roots lie above level D-1 (small
    trees) or split dims are explicit."""
    D = 3
    sizes = log_uniform_sizes(60, 1, 40000, 3, empty_every=11)
    off = offsets_of(sizes)
    if dist == "default":
        dist = "ties" if kind == "float" else "uniform"
    coords = forest_points(sizes, D, 5, dist=dist)
    lay = layout_for(kind, D)
    elems = make_elements(coords, lay)
    f = ck.Forest(off, lay)
    pts = torch.from_numpy(elems).cuda()
    wb = f.bounds_tensor()
    f.build(pts, wb)
    torch.cuda.synchronize()
    got = pts.cpu().numpy()
    gwb = wb.cpu().numpy()
    for t in range(len(sizes)):
        a, b = off[t], off[t + 1]
        if a == b:
            continue
        ref, rbox = oc.build_host(elems[a:b], lay, want_bounds=True)
        assert np.array_equal(got[a:b], ref), "tree %d (n=%d)" % (t, b - a)
        assert np.array_equal(gwb[t].view(rbox.dtype), rbox), "bounds of tree %d" % t
    f.close()


def _finish_blocks_at_largest_cap(sizes, D, explicit):
    """Finish workgroups of a forest if every tree took the largest LDS
    capacity (the forest plan's rule: 4096 for explicit dims with D > 4,
    8192 otherwise); 512 or more puts every tree on that capacity."""
    cap = 4096 if explicit and D > 4 else 8192
    total = 0
    for n in sizes:
        if n < 1:
            continue
        h = int(np.floor(np.log2(n - 1))) + 1 if n > 1 else 1
        total += 1 << max(0, h - int(np.log2(cap)))
    return total


@pytest.mark.parametrize("kind", ["explicit", "photon"])
@pytest.mark.parametrize("dist", ["gmm", "grid17", "coarse"])
def test_explicit_forest_with_many_finish_blocks(kind, dist):
    """An explicit-dim forest large enough (>= 512 finish workgroups) that
    every tree finishes from 8192-element subtrees: each tree equals the
    oracle's buildTree_host."""
    D = 3
    sizes = log_uniform_sizes(40, 40000, 300000, 61)
    assert _finish_blocks_at_largest_cap(sizes, D, True) >= 512
    off = offsets_of(sizes)
    coords = forest_points(sizes, D, 63, dist=dist)
    lay = layout_for(kind, D)
    elems = make_elements(coords, lay)
    f = ck.Forest(off, lay)
    pts = torch.from_numpy(elems).cuda()
    wb = f.bounds_tensor()
    f.build(pts, wb)
    torch.cuda.synchronize()
    got = pts.cpu().numpy()
    for t in range(len(sizes)):
        a, b = off[t], off[t + 1]
        ref, _ = oc.build_host(elems[a:b], lay, want_bounds=True)
        assert np.array_equal(got[a:b], ref), "tree %d (n=%d)" % (t, b - a)
    f.close()


def test_forest_equals_per_tree_builds_c4_shape():
    """512 trees, sizes log-uniform in [64, 262144]: the batched build and
    one cukd_build_tree per tree give identical arrays."""
    D = 3
    sizes = log_uniform_sizes(512, 64, 262144, 11)
    off = offsets_of(sizes)
    coords = forest_points(sizes, D, 13)
    lay = ck.float_points(D)
    orig = torch.from_numpy(coords).cuda()
    f = ck.Forest(off, lay)
    a = orig.clone()
    f.build(a)
    b = orig.clone()
    for t in range(len(sizes)):
        if sizes[t]:
            ck.build_tree(b[off[t]:off[t + 1]], lay)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    # and a few trees against the oracle
    for t in (0, 1, int(np.argmax(sizes))):
        ref, _ = oc.build_host(coords[off[t]:off[t + 1]], lay)
        assert np.array_equal(a[off[t]:off[t + 1]].cpu().numpy(), ref)
    f.close()


def test_forest_single_tree_and_all_empty():
    lay = ck.float_points(3)
    coords = datagen.uniform(70000, 3, 2)
    f = ck.Forest([0, 70000], lay)
    p = torch.from_numpy(coords).cuda()
    f.build(p)
    ref, _ = oc.build_host(coords, lay)
    torch.cuda.synchronize()
    assert np.array_equal(p.cpu().numpy(), ref)
    f.close()
    g = ck.Forest([0, 0, 0], lay)
    g.build(torch.empty((0, 3), dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    g.close()


def _segments(sizes, nq_per_tree, seed):
    """Segments in shuffled tree order, one tree listed twice."""
    T = len(sizes)
    order = np.argsort(datagen.splitmix64(seed, T))
    order = np.concatenate([order, order[:1]])
    seg_tree = order.astype(np.int32)
    counts = np.array([nq_per_tree[t] for t in seg_tree], np.int64)
    seg_off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return seg_off, seg_tree


@pytest.mark.parametrize("method", [ck.STACK_BASED, ck.STACK_FREE, ck.CCT])
def test_forest_queries_equal_per_tree_queries(method):
    D = 3
    sizes = log_uniform_sizes(80, 1, 30000, 21, empty_every=13)
    off = offsets_of(sizes)
    coords = forest_points(sizes, D, 23)
    lay = ck.float_points(D)
    f = ck.Forest(off, lay)
    pts = torch.from_numpy(coords).cuda()
    wb = f.bounds_tensor()
    f.build(pts, wb)
    nq_t = [int(max(1, s // 2)) for s in sizes]
    seg_off, seg_tree = _segments(sizes, nq_t, 5)
    # queries of segment s lie around tree seg_tree[s]
    qs = []
    for s, t in enumerate(seg_tree):
        n = int(seg_off[s + 1] - seg_off[s])
        q = datagen.uniform(n, D, 100 + s) * (1.0 + t % 3) * 1.1 + \
            datagen.uniform(1, D, 5 * 7 + int(t))[0] * 10.0
        qs.append(q.astype(np.float32))
    q = torch.from_numpy(np.concatenate(qs)).cuda()
    d_off = torch.from_numpy(seg_off).cuda()
    d_tree = torch.from_numpy(seg_tree).cuda()
    pts_h = pts.cpu().numpy()
    wb_h = wb.cpu().numpy()
    q_h = q.cpu().numpy()
    for k in (0, 1, 8, 20):
        if k == 0:
            ids, d2 = f.fcp(pts, q, d_off, d_tree, traversal=method, cutoff=0.5, world_bounds=wb)
        else:
            ids, d2, r2 = f.knn(pts, q, k, d_off, d_tree, traversal=method, world_bounds=wb)
        # the oracle on each segment's own tree (ids local to that tree)
        ids_h, d2_h = ids.cpu().numpy(), d2.cpu().numpy()
        for s, t in enumerate(seg_tree):
            a, b = int(seg_off[s]), int(seg_off[s + 1])
            tree_h = pts_h[off[t]:off[t + 1]]
            if k == 0:
                oid, od2 = oc.fcp(method, tree_h, lay, q_h[a:b], cutoff=0.5, world_bounds=wb_h[t])
            else:
                oid, od2, _ = oc.knn(method, oc.LIST_FIXED, k, tree_h, lay, q_h[a:b],
                                     world_bounds=wb_h[t])
            assert np.array_equal(ids_h[a:b], oid), ("oracle ids", k, s)
            assert np.array_equal(d2_h[a:b].view(np.uint32), od2.view(np.uint32)), ("oracle", k, s)
        for s, t in enumerate(seg_tree):
            a, b = int(seg_off[s]), int(seg_off[s + 1])
            tree = pts[off[t]:off[t + 1]]
            if k == 0:
                rid, rd2 = ck.fcp(tree, lay, q[a:b], traversal=method, cutoff=0.5,
                                  world_bounds=wb[t])
                assert torch.equal(ids[a:b], rid) and torch.equal(d2[a:b], rd2), (k, s)
            else:
                rid, rd2, rr = ck.knn(tree, lay, q[a:b], k, traversal=method, world_bounds=wb[t])
                assert torch.equal(ids[a:b], rid) and torch.equal(d2[a:b], rd2), (k, s)
                assert torch.equal(r2[a:b], rr)
    f.close()


def test_forest_build_and_query_in_a_hip_graph():
    """The forest build + kNN sequence captured once, replayed several
    times with eager work in between: every replay equals the eager run."""
    D, k = 3, 4
    sizes = log_uniform_sizes(200, 64, 50000, 31)
    off = offsets_of(sizes)
    coords = forest_points(sizes, D, 37)
    lay = ck.float_points(D)
    nq_t = [int(s // 2) for s in sizes]
    seg_tree = np.arange(len(sizes), dtype=np.int32)
    seg_off = np.concatenate([[0], np.cumsum(nq_t)]).astype(np.int32)
    q = torch.from_numpy(forest_points(nq_t, D, 37)).cuda()
    d_off = torch.from_numpy(seg_off).cuda()
    d_tree = torch.from_numpy(seg_tree).cuda()
    orig = torch.from_numpy(coords).cuda()
    f = ck.Forest(off, lay)
    # eager reference
    e_pts = orig.clone()
    f.build(e_pts)
    e_ids, e_d2, e_r = f.knn(e_pts, q, k, d_off, d_tree)
    torch.cuda.synchronize()
    # capture
    pts = orig.clone()
    ids = torch.empty_like(e_ids)
    d2 = torch.empty_like(e_d2)
    r = torch.empty_like(e_r)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up on the capture stream
        pts.copy_(orig)
        f.build(pts, stream=s)
        f.knn(pts, q, k, d_off, d_tree, stream=s, out=(ids, d2, r))
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        cs = torch.cuda.current_stream()
        pts.copy_(orig)
        f.build(pts, stream=cs)
        f.knn(pts, q, k, d_off, d_tree, stream=cs, out=(ids, d2, r))
    for rep in range(3):
        ids.fill_(-7)
        d2.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(pts, e_pts), rep
        assert torch.equal(ids, e_ids) and torch.equal(d2, e_d2) and torch.equal(r, e_r), rep
        _ = (orig * 2).sum().item()  # eager work between replays
    f.close()
