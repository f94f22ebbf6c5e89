"""Multi-round kNN on the GPU (cukd_knn_accumulate / cukd_knn_extract):
after every round the persistent lists are bit-identical to the oracle's
(knn.h:360-421 heap pushes in traversal order; samples/mpiHugeQuery.cu:75-103)."""
import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen

pytestmark = pytest.mark.gpu


def _shards(pts, cuts, lay):
    out, o = [], 0
    for a, b in zip(cuts[:-1], cuts[1:]):
        t, _ = oc.build_host(pts[a:b], lay)
        out.append((t, o, oc.compute_bounds(t, lay)))
        o += b - a
    return out


@pytest.mark.parametrize("method", [ck.STACK_BASED, ck.STACK_FREE, ck.CCT])
@pytest.mark.parametrize("k", [1, 4, 8, 16, 33, 64])
def test_rounds_match_oracle(method, k):
    D, nq = 3, 4000
    lay = ck.float_points(D)
    pts = datagen.gmm(60000, D, 5)
    q = datagen.uniform(nq, D, 6)
    shards = _shards(pts, [0, 9000, 30000, 30001, 60000], lay)
    ref = np.zeros((nq, k), np.uint64)
    lists = ck.knn_lists(nq, k)
    dq = torch.from_numpy(q).cuda()
    for r, (t, off, wb) in enumerate(shards):
        oc.knn_accumulate(method, k, t, lay, q, ref, init=(r == 0), cutoff=0.2, id_offset=off,
                          world_bounds=wb)
        dt = torch.from_numpy(t).cuda()
        dwb = torch.from_numpy(wb).cuda()
        ck.knn_accumulate(dt, lay, dq, lists, init=(r == 0), traversal=method, cutoff=0.2,
                          id_offset=off, world_bounds=dwb,
                          flags=ck.QUERY_SORT if r % 2 else 0)
        torch.cuda.synchronize()
        assert np.array_equal(lists.cpu().numpy().view(np.uint64), ref), "round %d" % r
    ids, d2, rad = ck.knn_extract(lists)
    rid, rd2, rrad = oc.knn_extract(ref)
    torch.cuda.synchronize()
    assert np.array_equal(ids.cpu().numpy(), rid)
    assert np.array_equal(d2.cpu().numpy().view(np.uint32), rd2.view(np.uint32))
    assert np.array_equal(rad.cpu().numpy().view(np.uint32), rrad.view(np.uint32))


def test_single_round_equals_heap_list_query():
    D, nq, k = 4, 20000, 8
    lay = ck.float_points(D)
    pts = datagen.uniform(100000, D, 7)
    tree, _ = oc.build_host(pts, lay)
    dt = torch.from_numpy(tree).cuda()
    dq = torch.from_numpy(datagen.uniform(nq, D, 8)).cuda()
    ids, d2, _ = ck.knn(dt, lay, dq, k, traversal=ck.STACK_FREE, list_kind=ck.LIST_HEAP)
    lists = ck.knn_lists(nq, k)
    ck.knn_accumulate(dt, lay, dq, lists, init=True)
    torch.cuda.synchronize()
    keys = (d2.cpu().numpy().view(np.uint32).astype(np.uint64) << np.uint64(32)) | \
        ids.cpu().numpy().view(np.uint32).astype(np.uint64)
    assert np.array_equal(lists.cpu().numpy().view(np.uint64), keys)


def test_empty_tree_round_and_infinite_radius():
    D, nq, k = 3, 1000, 8
    lay = ck.float_points(D)
    q = datagen.uniform(nq, D, 9)
    dq = torch.from_numpy(q).cuda()
    lists = ck.knn_lists(nq, k)
    empty = torch.empty((0, D), dtype=torch.float32, device="cuda")
    ck.knn_accumulate(empty, lay, dq, lists, init=True)
    ids, d2, rad = ck.knn_extract(lists)
    torch.cuda.synchronize()
    assert torch.all(ids == -1) and torch.all(torch.isinf(rad))
    tree, _ = oc.build_host(datagen.uniform(3, D, 10), lay)  # fewer points than k
    ref = np.zeros((nq, k), np.uint64)
    oc.knn_accumulate(oc.STACK_FREE, k, tree, lay, q, ref, init=True)
    ck.knn_accumulate(torch.from_numpy(tree).cuda(), lay, dq, lists, init=False)
    torch.cuda.synchronize()
    assert np.array_equal(lists.cpu().numpy().view(np.uint64), ref)
