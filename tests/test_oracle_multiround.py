"""CPU checks of the oracle's multi-round kNN (persistent FlexHeap lists,
knn.h:266-287, 360-421; runQuery / extractFinalResult,
samples/mpiHugeQuery.cu:75-103).

Pinned by brute force: after rounds over the trees of a sharded point set,
each list holds the k smallest (dist2, global id) keys over the union of
the shards, whatever order the shards come in; one round from an
initialised list is the reference's HeapCandidateList query."""
import numpy as np
import pytest

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen


def _brute_d2(pts, q):
    d = pts[None, :, :] - q[:, None, :]
    s = d[..., 0] * d[..., 0]
    for i in range(1, pts.shape[1]):
        s = s + d[..., i] * d[..., i]
    return s


def _keys(d2, ids):
    return (d2.view(np.uint32).astype(np.uint64) << np.uint64(32)) | ids.view(np.uint32).astype(np.uint64)


def _sharded_trees(pts, cuts, lay):
    trees, offs, o = [], [], 0
    for a, b in zip(cuts[:-1], cuts[1:]):
        t, _ = oc.build_host(pts[a:b], lay)
        trees.append(t)
        offs.append(o)
        o += b - a
    return trees, offs


@pytest.mark.parametrize("method", [oc.STACK_BASED, oc.STACK_FREE, oc.CCT])
@pytest.mark.parametrize("k", [1, 4, 8, 20])
@pytest.mark.parametrize("cutoff", [np.inf, 0.08])
def test_rounds_equal_bruteforce_over_union(method, k, cutoff):
    D, nq = 3, 200
    pts = datagen.uniform(5000, D, 11)
    q = datagen.uniform(nq, D, 12)
    lay = ck.float_points(D)
    trees, offs = _sharded_trees(pts, [0, 700, 2900, 3000, 5000], lay)
    lists = np.zeros((nq, k), np.uint64)
    for r, (t, o) in enumerate(zip(trees, offs)):
        wb = oc.compute_bounds(t, lay)
        oc.knn_accumulate(method, k, t, lay, q, lists, init=(r == 0), cutoff=cutoff,
                          id_offset=o, world_bounds=wb)
    union = np.concatenate(trees)
    full = _brute_d2(union, q)
    gid = np.broadcast_to(np.arange(len(union), dtype=np.int32), full.shape)
    keys = np.sort(_keys(full, np.ascontiguousarray(gid)), 1)
    r2 = np.float32(cutoff) * np.float32(cutoff)
    empty = _keys(np.full(k, r2, np.float32), np.full(k, -1, np.int32))
    want = np.sort(np.concatenate([keys[:, :k], np.broadcast_to(empty, (nq, k))], 1), 1)[:, :k]
    assert np.array_equal(np.sort(lists, 1), want)


def test_round_order_does_not_change_the_set():
    D, nq, k = 3, 150, 8
    pts = datagen.gmm(6000, D, 3)
    q = datagen.uniform(nq, D, 4)
    lay = ck.float_points(D)
    trees, offs = _sharded_trees(pts, [0, 1000, 4000, 6000], lay)
    sets = []
    for order in ([0, 1, 2], [2, 0, 1], [1, 2, 0]):
        lists = np.zeros((nq, k), np.uint64)
        for r, i in enumerate(order):
            oc.knn_accumulate(oc.STACK_FREE, k, trees[i], lay, q, lists, init=(r == 0),
                              id_offset=offs[i])
        sets.append(np.sort(lists, 1))
    assert np.array_equal(sets[0], sets[1]) and np.array_equal(sets[0], sets[2])


@pytest.mark.parametrize("method", [oc.STACK_BASED, oc.STACK_FREE, oc.CCT])
def test_single_round_is_the_heap_list_query(method):
    D, nq, k = 3, 300, 16
    pts = datagen.uniform(4000, D, 21)
    q = datagen.uniform(nq, D, 22)
    lay = ck.float_points(D)
    tree, _ = oc.build_host(pts, lay)
    wb = oc.compute_bounds(pts, lay)
    ids, d2, _ = oc.knn(method, oc.LIST_HEAP, k, tree, lay, q, cutoff=0.1, world_bounds=wb)
    lists = np.zeros((nq, k), np.uint64)
    oc.knn_accumulate(method, k, tree, lay, q, lists, init=True, cutoff=0.1, world_bounds=wb)
    assert np.array_equal(lists, _keys(d2, ids))


def test_extract():
    D, nq, k = 3, 100, 8
    lay = ck.float_points(D)
    tree, _ = oc.build_host(datagen.uniform(5, D, 1), lay)  # fewer points than k
    q = datagen.uniform(nq, D, 2)
    lists = np.zeros((nq, k), np.uint64)
    oc.knn_accumulate(oc.STACK_BASED, k, tree, lay, q, lists, init=True)
    ids, d2, rad = oc.knn_extract(lists)
    assert np.all(np.isinf(rad))  # extractFinalResult keeps infinity
    assert np.array_equal(_keys(d2, ids), np.sort(lists, 1))
    assert np.all(ids[:, 5:] == -1) and np.all(ids[:, :5] >= 0)
    tree2, _ = oc.build_host(datagen.uniform(3000, D, 3), lay)
    oc.knn_accumulate(oc.STACK_BASED, k, tree2, lay, q, lists, init=False, id_offset=5)
    ids, d2, rad = oc.knn_extract(lists)
    m = np.sort(lists, 1)[:, -1] >> np.uint64(32)
    assert np.array_equal(rad, np.sqrt(m.astype(np.uint32).view(np.float32)))
    assert np.all(np.diff(d2, axis=1) >= 0)
