"""Pins the CPU oracle (oracle/cukd_oracle.c) without the reference binary.

Each check is independent of the oracle's own code path:
  * tree-index closed forms  vs  brute-force counting on explicit trees;
  * the host builder         vs  (a) a recursive median-partition builder
                                 on tie-free inputs (the reference's own
                                 cross-builder check,
                                 testing/testBuildersSameResult.cu), and
                                 (b) a Python stable-sort replay of the
                                 per-level (tag, coord) sort, on tie-heavy
                                 inputs;
  * queries                  vs  brute force with the same FP order;
  * the issue5 known answer  (testing/issue5.cu:77-107).
"""
import numpy as np
import pytest

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.helpers import (check_kdtree_valid, coords_of, dims_of, distribution,
                           layout_for, make_elements)


# ---------------------------------------------------------------------------
# tree index math (helpers.h:76-258)
# ---------------------------------------------------------------------------

def _brute_subtree_size(node, n):
    cnt, frontier = 0, [node]
    while frontier:
        nxt = []
        for x in frontier:
            if x < n:
                cnt += 1
                nxt += [2 * x + 1, 2 * x + 2]
        frontier = nxt
    return cnt


@pytest.mark.parametrize("n", list(range(1, 130)) + [255, 256, 257, 1000, 1023, 1024, 1025, 4095])
def test_closed_forms_vs_bruteforce(n):
    levels = oc.num_levels_for(n)
    assert (1 << (levels - 1)) <= n < (1 << levels)
    for node in range(min(n, 700)):
        assert oc.subtree_size(node, n) == _brute_subtree_size(node, n)
        assert oc.level_of(node) == int(np.floor(np.log2(node + 1)))
    # segment layout of step L: settled prefix, then level-L subtrees in order
    for step in range(levels - 1):
        first = (1 << step) - 1
        begin = first
        for sub in range(first, min(2 * first + 1, n)):
            assert oc.segment_begin(step, n, sub) == begin
            left = 2 * sub + 1
            lsz = _brute_subtree_size(left, n) if left < n else 0
            assert oc.pivot_pos(step, n, sub) == begin + lsz
            begin += _brute_subtree_size(sub, n)


def test_widest_dim_tie_rule():
    lay = ck.float_points(3)
    assert oc.widest_dim(np.array([0, 0, 0, 1, 1, 1], np.float32), lay) == 2
    assert oc.widest_dim(np.array([0, 0, 0, 2, 1, 2], np.float32), lay) == 2
    assert oc.widest_dim(np.array([0, 0, 0, 3, 1, 2], np.float32), lay) == 0
    assert oc.widest_dim(np.array([0, 0, 0, 0, 0, 0], np.float32), lay) == 2
    lay_i = ck.int_points(2)
    assert oc.widest_dim(np.array([0, 0, 5, 5], np.int32), lay_i) == 1


# ---------------------------------------------------------------------------
# builder: independent restatements
# ---------------------------------------------------------------------------

def _widest(lo, hi):
    best, wb = 0, 0
    for d in range(len(lo)):
        w = hi[d] - lo[d]
        if w < wb:
            continue
        wb, best = w, d
    return best


def _partition_builder(coords, explicit, world=None, composite=False):
    """Recursive median selection (ranks under the composite key
    (c[d_L], c[d_{L-1}], ..., input index) when composite, else by the
    split coordinate alone, valid for tie-free inputs)."""
    n, D = coords.shape
    order = np.empty(n, np.int64)
    dims = np.zeros(n, np.int64)
    stack = [(0, np.arange(n), [], None if world is None else (world[0].copy(), world[1].copy()))]
    while stack:
        node, members, chain, box = stack.pop()
        if node >= n:
            continue
        lvl = int(np.floor(np.log2(node + 1)))
        d = _widest(*box) if explicit else lvl % D
        chain2 = [d] + [c for c in chain if c != d]
        if composite:
            keys = tuple(coords[members, c] for c in reversed(chain2[1:])) + (coords[members, d],)
            srt = members[np.lexsort((members,) + keys)]
        else:
            srt = members[np.argsort(coords[members, d], kind="stable")]
        left = oc.subtree_size(2 * node + 1, n) if 2 * node + 1 < n else 0
        piv = srt[left]
        order[node] = piv
        dims[node] = d
        if explicit:
            lb = (box[0].copy(), box[1].copy()); lb[1][d] = coords[piv, d]
            rb = (box[0].copy(), box[1].copy()); rb[0][d] = coords[piv, d]
        else:
            lb = rb = None
        stack.append((2 * node + 1, srt[:left], chain2, lb))
        stack.append((2 * node + 2, srt[left + 1:], chain2, rb))
    return order, dims


def _stable_replay_builder(coords, explicit, world):
    """The per-level (tag, coord) stable sort of buildTree_host replayed with
    Python's (stable) sort and brute-force subtree sizes."""
    n, D = coords.shape
    idx = list(range(n))
    tags = [0] * n
    dims = [(_widest(*world) if explicit else 0)] * n
    levels = oc.num_levels_for(n)

    def seg_begin(step, sub):
        first = (1 << step) - 1
        return first + sum(_brute_subtree_size(s, n) for s in range(first, sub))

    for level in range(levels - 1):
        pairs = sorted(range(n), key=lambda p: (tags[p], coords[idx[p], dims[idx[p]] if explicit else level % D]))
        idx = [idx[p] for p in pairs]
        tags = [tags[p] for p in pairs]
        settled = (1 << level) - 1
        new_tags = list(tags)
        new_dims = list(dims)
        for gid in range(settled, n):
            sub = tags[gid]
            pivot = seg_begin(level, sub) + (_brute_subtree_size(2 * sub + 1, n) if 2 * sub + 1 < n else 0)
            if gid == pivot:
                continue
            child = 2 * sub + 1 if gid < pivot else 2 * sub + 2
            new_tags[gid] = child
            if explicit:
                # segment domain: walk up from the segment's node through the
                # settled prefix, then clip at this level's pivot
                lo, hi = world[0].copy(), world[1].copy()
                cur = sub
                while cur > 0:
                    par = (cur + 1) // 2 - 1
                    pd = dims[idx[par]]
                    s = coords[idx[par], pd]
                    if cur & 1:
                        hi[pd] = min(hi[pd], s)
                    else:
                        lo[pd] = max(lo[pd], s)
                    cur = par
                pdim = dims[idx[pivot]]
                if gid < pivot:
                    hi[pdim] = coords[idx[pivot], pdim]
                else:
                    lo[pdim] = coords[idx[pivot], pdim]
                new_dims[idx[gid]] = _widest(lo, hi)
        tags = new_tags
        dims = new_dims
    pairs = sorted(range(n), key=lambda p: tags[p])
    idx = [idx[p] for p in pairs]
    return np.array(idx), np.array([dims[i] for i in idx])


def _oracle_build(coords, kind):
    lay = layout_for(kind, coords.shape[1])
    elems = make_elements(coords, lay)
    tree, box = oc.build_host(elems, lay, want_bounds=True)
    return lay, elems, tree, box


def _order_of(tree, elems, D):
    # the payload word (right after the coordinates) identifies the element
    key = {int(e[D]): i for i, e in enumerate(elems)}
    assert len(key) == len(elems)
    return np.array([key[int(t[D])] for t in tree])


@pytest.mark.parametrize("D", [2, 3, 4, 8])
@pytest.mark.parametrize("kind", ["float", "int", "payload", "explicit", "explicit_int"])
@pytest.mark.parametrize("n", [4, 10, 1000])
def test_host_builder_tie_free_equals_partition(D, kind, n):
    scalar_float = kind in ("float", "payload", "explicit")
    c = datagen.unique_ints(n, D, seed=n * 31 + D)
    coords = c.astype(np.float32) if scalar_float else c
    lay, elems, tree, box = _oracle_build(coords, kind)
    explicit = lay.dim_offset >= 0
    world = (coords.min(0), coords.max(0))
    order, dims = _partition_builder(coords, explicit, world)
    t_coords = coords_of(tree, lay)
    assert np.array_equal(t_coords, coords[order])
    if explicit:
        tdims = dims_of(tree, lay)
        assert np.array_equal(tdims, dims)
        np.testing.assert_array_equal(box, np.concatenate(world))
    check_kdtree_valid(t_coords, dims if explicit else
                       (np.floor(np.log2(np.arange(n) + 1)).astype(np.int64) % D), n)


@pytest.mark.parametrize("dist", ["ties", "dups", "plane", "signed_zero"])
@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("explicit", [False, True])
def test_host_builder_ties_equal_stable_replay_and_composite_key(dist, D, explicit):
    n = 157
    coords = distribution(dist, n, D, seed=5)
    kind = "explicit" if explicit else "payload"
    lay, elems, tree, box = _oracle_build(coords, kind)
    world = (coords.min(0), coords.max(0))
    got = _order_of(tree, elems, D)
    replay, rdims = _stable_replay_builder(coords, explicit, world)
    assert np.array_equal(got, replay)
    comp, cdims = _partition_builder(coords, explicit, world, composite=True)
    assert np.array_equal(got, comp)
    if explicit:
        assert np.array_equal(dims_of(tree, lay), rdims)


@pytest.mark.parametrize("dist", ["uniform", "ties", "gmm", "plane"])
def test_host_builder_composite_key_larger(dist):
    n, D = 3000, 3
    coords = distribution(dist, n, D, seed=11)
    lay, elems, tree, box = _oracle_build(coords, "payload")
    comp, _ = _partition_builder(coords, False, composite=True)
    assert np.array_equal(_order_of(tree, elems, D), comp)


def test_host_builder_empty_and_single():
    lay = ck.float_points(3)
    t, _ = oc.build_host(np.zeros((0, 3), np.float32), lay)
    assert t.shape == (0, 3)
    p = np.array([[1.0, 2.0, 3.0]], np.float32)
    t, _ = oc.build_host(p, lay)
    assert np.array_equal(t, p)


def test_hash_matches_reference_formula():
    a = np.array([1, 2, 3, 0xFFFFFFFF], np.uint32)
    h = 0
    for w in a:
        h = ((h * 17) & (2 ** 64 - 1)) ^ int(w)
    assert oc.hash_words(a) == h


# ---------------------------------------------------------------------------
# queries
# ---------------------------------------------------------------------------

def _brute_d2(tree_coords, q):
    d = tree_coords[None, :, :] - q[:, None, :]
    s = d[..., 0] * d[..., 0]
    for i in range(1, tree_coords.shape[1]):
        s = s + d[..., i] * d[..., i]
    return s


def test_issue5_known_answer():
    pts = datagen.msvc_rand_points(90167, 33)
    # issue5.cu:80-84
    expect = np.array([[-99.1088562, -87.9879150, 27.7626877],
                       [-38.3892326, 31.5713978, -37.0891457],
                       [-22.0435200, -92.5473785, 89.4833221],
                       [48.5274811, -94.0671997, 80.3888092],
                       [-33.9030113, 34.4157219, 95.2085953]], np.float32)
    np.testing.assert_allclose(pts[:5], expect, rtol=0, atol=2e-5)
    lay = ck.float_points(3)
    tree, _ = oc.build_host(pts, lay)
    wb = oc.compute_bounds(pts, lay)
    q = np.array([[-98.4496613, 76.9219055, 25.8888512]], np.float32)
    for m in (oc.STACK_BASED, oc.STACK_FREE, oc.CCT):
        ids, d2 = oc.fcp(m, tree, lay, q, cutoff=5.0, world_bounds=wb)
        assert d2.view(np.uint32)[0] == 0x40362F30  # float(2.8466301f)
        assert np.float32(2.8466301) == d2[0]
        assert _brute_d2(tree, q)[0, ids[0]] == d2[0]


@pytest.mark.parametrize("D", [2, 3, 4, 8])
@pytest.mark.parametrize("dist", ["uniform", "gmm", "ties"])
def test_queries_vs_bruteforce(D, dist):
    n, nq = 4000, 300
    coords = distribution(dist, n, D, seed=3)
    lay = ck.float_points(D)
    tree, _ = oc.build_host(coords, lay)
    wb = oc.compute_bounds(coords, lay)
    q = datagen.uniform(nq, D, 17)
    full = _brute_d2(tree, q)
    for m in (oc.STACK_BASED, oc.STACK_FREE, oc.CCT):
        for cutoff in (np.inf, 0.05):
            ids, d2 = oc.fcp(m, tree, lay, q, cutoff=cutoff, world_bounds=wb)
            r2 = np.float32(cutoff) * np.float32(cutoff)
            bmin = full.min(1)
            has = bmin < r2
            assert np.array_equal(ids >= 0, has)
            assert np.array_equal(d2[has], bmin[has])
            assert np.array_equal(full[np.nonzero(has)[0], ids[has]], bmin[has])
            for lk in (oc.LIST_FIXED, oc.LIST_HEAP):
                for k in (1, 5, 8):
                    kid, kd2, ret = oc.knn(m, lk, k, tree, lay, q, cutoff=cutoff, world_bounds=wb)
                    srt = np.sort(full, 1)[:, :k]
                    cand = np.where(srt <= r2, srt, r2)
                    got = np.sort(kd2, 1)
                    assert np.array_equal(got, cand)
                    kth = cand[:, -1]
                    assert np.array_equal(ret, kth)
                    valid = kid >= 0
                    rows = np.nonzero(valid)
                    assert np.array_equal(full[rows[0], kid[valid]], kd2[valid])
                    if lk == oc.LIST_FIXED:
                        assert np.all(np.diff(kd2, axis=1) >= 0)


def test_knn_fixed_and_heap_hold_same_set():
    n, D = 5000, 3
    coords = distribution("ties", n, D, seed=9)
    lay = ck.float_points(D)
    tree, _ = oc.build_host(coords, lay)
    q = datagen.uniform(200, D, 4) * 7
    a = oc.knn(oc.STACK_BASED, oc.LIST_FIXED, 16, tree, lay, q)
    b = oc.knn(oc.STACK_BASED, oc.LIST_HEAP, 16, tree, lay, q)
    ka = np.sort(a[1].view(np.uint32).astype(np.uint64) << np.uint64(32) | a[0].view(np.uint32), 1)
    kb = np.sort(b[1].view(np.uint32).astype(np.uint64) << np.uint64(32) | b[0].view(np.uint32), 1)
    assert np.array_equal(ka, kb)
    assert np.array_equal(a[2], b[2])


def test_explicit_dim_tree_queries():
    n, D = 3000, 3
    coords = datagen.gmm(n, D, 2)
    lay = layout_for("explicit", D)
    elems = make_elements(coords, lay)
    tree, box = oc.build_host(elems, lay)
    tc = coords_of(tree, lay)
    check_kdtree_valid(tc, dims_of(tree, lay), n)
    q = datagen.uniform(200, D, 8)
    full = _brute_d2(tc, q)
    for m in (oc.STACK_BASED, oc.STACK_FREE, oc.CCT):
        ids, d2 = oc.fcp(m, tree, lay, q, world_bounds=box)
        assert np.array_equal(d2, full.min(1))


# ---------------------------------------------------------------------------
# spatial k-d tree
# ---------------------------------------------------------------------------

def _check_spatial(nodes, prims, bounds, coords, leaf):
    n, D = coords.shape
    seen = np.zeros(n, np.int64)
    stack = [(0, bounds[:D].copy(), bounds[D:].copy(), 0)]
    while stack:
        nid, lo, hi, depth = stack.pop()
        nd = nodes[nid]
        if nd["count"] > 0:
            assert nd["count"] <= leaf
            ids = prims[nd["offset"]:nd["offset"] + nd["count"]]
            assert np.all(np.diff(ids.astype(np.int64)) > 0)  # ascending primID
            seen[ids] += 1
            p = coords[ids]
            assert np.all(p >= lo) and np.all(p <= hi)
            continue
        d, s = nd["dim"], nd["pos"]
        l_hi = hi.copy(); l_hi[d] = s
        r_lo = lo.copy(); r_lo[d] = s
        stack.append((nd["offset"], lo, l_hi, depth + 1))
        stack.append((nd["offset"] + 1, r_lo, hi, depth + 1))
    assert np.all(seen == 1)


@pytest.mark.parametrize("dist", ["uniform", "gmm"])
@pytest.mark.parametrize("D", [2, 3, 4])
def test_spatial_build_and_queries(dist, D):
    n = 5000
    coords = distribution(dist, n, D, seed=21)
    lay = ck.float_points(D)
    nodes, prims, bounds = oc.spatial_build(coords, lay)
    np.testing.assert_array_equal(bounds, np.concatenate([coords.min(0), coords.max(0)]))
    _check_spatial(nodes, prims, bounds, coords, 8)
    q = datagen.uniform(300, D, 5)
    full = _brute_d2(coords, q)
    for m in (oc.STACK_BASED, oc.CCT):
        ids, d2 = oc.spatial_fcp(m, nodes, prims, bounds, coords, lay, q)
        assert np.array_equal(d2, full.min(1))
        assert np.array_equal(full[np.arange(len(q)), ids], d2)
        for lk in (oc.LIST_FIXED, oc.LIST_HEAP):
            kid, kd2, ret = oc.spatial_knn(m, lk, 16, nodes, prims, bounds, coords, lay, q)
            assert np.array_equal(np.sort(kd2, 1), np.sort(full, 1)[:, :16])


def test_spatial_degenerate_duplicates():
    n, D = 1000, 3
    coords = np.repeat(datagen.uniform(10, D, 3), 100, axis=0)
    lay = ck.float_points(D)
    nodes, prims, bounds = oc.spatial_build(coords, lay)
    _check_spatial(nodes, prims, bounds, coords, 8)


def test_spatial_empty():
    lay = ck.float_points(3)
    nodes, prims, bounds = oc.spatial_build(np.zeros((0, 3), np.float32), lay)
    assert len(nodes) == 2 and nodes[0]["count"] == 0
    ids, d2 = oc.spatial_fcp(oc.STACK_BASED, nodes, prims, bounds,
                             np.zeros((1, 3), np.float32), lay, datagen.uniform(4, 3, 1))
    assert np.all(ids == -1)


def test_spatial_int_build_is_the_float_build_of_converted_coordinates():
    """The reference builds int points through float: AtomicBox encodes
    float(coord), the side test compares float(coord) >= pos, and Node::pos
    / the bounds are scalar_t, so they hold the float value truncated to
    int (spatial-kdtree.h:172-188, 427-429, 489).  With coordinates below
    2^24 the conversion is exact, so the int tree must equal the float tree
    over the same values in everything but the truncated pos and bounds."""
    n = 20000
    c = (datagen.splitmix64(5, n * 3) >> np.uint64(44)).astype(np.int32).reshape(n, 3)
    c[:200] = c[200:400]                        # exact duplicates
    c[400:600, 1] = 77                          # a degenerate axis
    ni, pi, bi = oc.spatial_build(c, ck.int_points(3))
    nf, pf, bf = oc.spatial_build(c.astype(np.float32), ck.float_points(3))
    assert np.array_equal(pi, pf)
    for f in ("offset", "count", "dim"):
        assert np.array_equal(ni[f], nf[f])
    ipos = ni.view(np.uint8).reshape(-1, 12)[:, :4].copy().view(np.int32).ravel()
    assert np.array_equal(ipos, np.trunc(nf["pos"]).astype(np.int32))
    assert bi.dtype == np.int32 and np.array_equal(bi, np.trunc(bf).astype(np.int32))
    # some split lands between integers, so the truncation is exercised
    inner = nf["count"] == 0
    assert np.any(nf["pos"][inner] != np.trunc(nf["pos"][inner]))
