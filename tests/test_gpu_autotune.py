"""The query autotuner (CUKD_QUERY_AUTOTUNE) and Morton-ordered forest
queries: every launch configuration gives the plain launch's results."""
import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.test_gpu_forest import forest_points, log_uniform_sizes, offsets_of

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", [ck.STACK_BASED, ck.STACK_FREE, ck.CCT])
def test_autotuned_queries_equal_oracle(method):
    D = 3
    lay = ck.float_points(D)
    pts = datagen.uniform(200000, D, 3)
    tree, _ = oc.build_host(pts, lay)
    wb = oc.compute_bounds(pts, lay)
    q = datagen.uniform(30000, D, 4)
    dt, dq, dwb = (torch.from_numpy(x).cuda() for x in (tree, q, wb))
    rid, rd2, rr = oc.knn(method, oc.LIST_FIXED, 8, tree, lay, q, world_bounds=wb)
    fid, fd2 = oc.fcp(method, tree, lay, q, world_bounds=wb)
    for rep in range(2):  # first call tunes, second uses the cached choice
        ids, d2, r = ck.knn(dt, lay, dq, 8, traversal=method, world_bounds=dwb,
                            flags=ck.QUERY_AUTOTUNE)
        gi, gd = ck.fcp(dt, lay, dq, traversal=method, world_bounds=dwb,
                        flags=ck.QUERY_AUTOTUNE | ck.QUERY_SORT)
        torch.cuda.synchronize()
        assert np.array_equal(ids.cpu().numpy(), rid) and np.array_equal(d2.cpu().numpy(), rd2)
        assert np.array_equal(r.cpu().numpy(), rr)
        assert np.array_equal(gi.cpu().numpy(), fid) and np.array_equal(gd.cpu().numpy(), fd2)


def test_autotune_keeps_statistics_exact():
    D = 3
    lay = ck.float_points(D)
    tree, _ = oc.build_host(datagen.uniform(50000, D, 5), lay)
    q = datagen.uniform(7000, D, 6)
    _, _, _, vis = oc.knn(oc.STACK_BASED, oc.LIST_FIXED, 4, tree, lay, q, want_visits=True)
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    ck.set_traversal_stats(counter)
    try:
        ck.knn(torch.from_numpy(tree).cuda(), lay, torch.from_numpy(q).cuda(), 4,
               flags=ck.QUERY_AUTOTUNE)
        torch.cuda.synchronize()
    finally:
        ck.set_traversal_stats(None)
    assert int(counter.item()) == int(vis.sum())


def test_forest_sorted_and_autotuned_queries():
    D, k = 3, 4
    sizes = log_uniform_sizes(100, 64, 60000, 61)
    off = offsets_of(sizes)
    lay = ck.float_points(D)
    pts = torch.from_numpy(forest_points(sizes, D, 63)).cuda()
    f = ck.Forest(off, lay)
    f.build(pts)
    nq_t = sizes // 2
    seg_off = torch.from_numpy(np.concatenate([[0], np.cumsum(nq_t)]).astype(np.int32)).cuda()
    seg_tree = torch.arange(len(sizes), dtype=torch.int32, device="cuda")
    q = torch.from_numpy(forest_points(nq_t, D, 63)).cuda()
    base = f.knn(pts, q, k, seg_off, seg_tree)
    f.reserve_queries(len(q))
    for flags in (ck.QUERY_SORT, ck.QUERY_AUTOTUNE, ck.QUERY_AUTOTUNE):
        got = f.knn(pts, q, k, seg_off, seg_tree, flags=flags)
        torch.cuda.synchronize()
        for a, b in zip(got, base):
            assert torch.equal(a, b), flags
    f.close()




MORTON_BITS = 15  # csrc/qorder.hip


def test_query_order_is_a_morton_permutation():
    """cukd_query_order returns a permutation whose queries have
    non-decreasing Morton cells (50000 queries: several counting tiles),
    and launching in that order changes no result."""
    D = 3
    q = datagen.gmm(50000, D, 11)
    dq = torch.from_numpy(q).cuda()
    order = ck.query_order(dq).cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(len(q)))
    lo, hi = q.min(0), q.max(0)  # float32, as on the device
    t = (q - lo) / (hi - lo)
    b = MORTON_BITS // D  # bits per dim
    cell = np.clip((t * np.float32(1 << b)).astype(np.int64), 0, (1 << b) - 1)
    code = np.zeros(len(q), np.int64)
    for j in range(b):
        for d in range(D):
            code |= ((cell[:, d] >> j) & 1) << (j * D + d)
    assert np.all(np.diff(code[order]) >= 0)
    lay = ck.float_points(D)
    tree, _ = oc.build_host(datagen.uniform(20000, D, 12), lay)
    dt = torch.from_numpy(tree).cuda()
    a = ck.knn(dt, lay, dq, 4)
    b = ck.knn(dt, lay, dq[torch.from_numpy(order).cuda().long()].contiguous(), 4)
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    for x, y in zip(a, b):
        assert np.array_equal(x.cpu().numpy(), y.cpu().numpy()[inv])
