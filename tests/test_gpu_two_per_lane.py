"""Two queries per lane (the default stack-based launch for FCP and
FixedCandidateList k = 1/4/8/16) and one per lane (CUKD_QUERY_ONE_PER_LANE)
both return the oracle's ids, dist2 bits and return values exactly
(traverse-default-stack-based.h:24-108), in the plain and Morton launch
orders, on single trees and forests, and count the same node visits;
lists and trees the two-query kernel does not serve run one per lane."""
import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.helpers import distribution
from tests.test_gpu_forest import forest_points, log_uniform_sizes, offsets_of
from tests.test_gpu_query import _setup, check_fcp, check_knn

pytestmark = pytest.mark.gpu

TWO = 0  # the default launch
ONE = ck.QUERY_ONE_PER_LANE
SB = (ck.STACK_BASED,)


@pytest.mark.parametrize("dist", ["uniform", "gmm", "ties"])
@pytest.mark.parametrize("flags", [TWO, TWO | ck.QUERY_SORT, ONE, ONE | ck.QUERY_SORT])
def test_two_per_lane_equals_oracle(dist, flags):
    s = _setup(distribution(dist, 60000, 3, seed=31))
    q = datagen.uniform(20001, 3, 32)  # odd: the last lane has one query
    for cutoff in (float("inf"), 0.02):
        check_fcp(*s, q, cutoff, methods=SB, flags=flags)
        for k in (1, 4, 8, 16):
            check_knn(*s, q, k, cutoff, lists=(ck.LIST_FIXED,), methods=SB, flags=flags)


@pytest.mark.parametrize("D", [1, 2, 4, 8])
def test_two_per_lane_other_dims(D):
    s = _setup(distribution("gmm", 20000, D, seed=33))
    q = datagen.uniform(5000, D, 34)
    check_fcp(*s, q, float("inf"), methods=SB, flags=TWO)
    check_knn(*s, q, 8, float("inf"), lists=(ck.LIST_FIXED,), methods=SB, flags=TWO)


@pytest.mark.parametrize("nq", [1, 2, 63, 64, 65, 127, 128, 129, 257])
def test_two_per_lane_ragged_batches(nq):
    s = _setup(distribution("uniform", 3000, 3, seed=35))
    q = datagen.uniform(nq, 3, 36)
    check_knn(*s, q, 4, float("inf"), lists=(ck.LIST_FIXED,), methods=SB, flags=TWO)
    check_fcp(*s, q, float("inf"), methods=SB, flags=TWO | ck.QUERY_SORT)


@pytest.mark.parametrize("n", [0, 1, 2, 3, 7])
def test_two_per_lane_tiny_trees(n):
    lay = ck.float_points(3)
    q = datagen.uniform(300, 3, 37)
    if n == 0:
        dt = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
        ids, d2, r = ck.knn(dt, lay, torch.from_numpy(q).cuda(), 4, flags=TWO)
        torch.cuda.synchronize()
        assert np.all(ids.cpu().numpy() == -1)
        assert np.all(np.isinf(r.cpu().numpy()))
        return
    s = _setup(datagen.uniform(n, 3, 38))
    check_knn(*s, q, 4, float("inf"), lists=(ck.LIST_FIXED,), methods=SB, flags=TWO)
    check_fcp(*s, q, 0.3, methods=SB, flags=TWO)


def test_one_per_lane_where_two_do_not_apply():
    """Heap lists, run-time k, other traversals and explicit-dim trees run
    one query per lane and still match the oracle."""
    s = _setup(distribution("gmm", 20000, 3, seed=39))
    q = datagen.uniform(3000, 3, 40)
    check_knn(*s, q, 5, float("inf"), flags=TWO)  # run-time k, both lists, all methods
    check_knn(*s, q, 8, float("inf"), lists=(ck.LIST_HEAP,), flags=TWO)
    check_fcp(*s, q, float("inf"), flags=TWO)
    e = _setup(distribution("gmm", 20000, 3, seed=41), "explicit")
    check_knn(*e, q, 8, float("inf"), lists=(ck.LIST_FIXED,), methods=SB, flags=TWO)
    check_fcp(*e, q, float("inf"), methods=SB, flags=TWO)


@pytest.mark.parametrize("flags", [TWO, ONE])
def test_two_per_lane_statistics_exact(flags):
    lay = ck.float_points(3)
    tree, _ = oc.build_host(datagen.uniform(50000, 3, 42), lay)
    q = datagen.uniform(7001, 3, 43)
    _, _, _, vis = oc.knn(oc.STACK_BASED, oc.LIST_FIXED, 8, tree, lay, q, want_visits=True)
    counter = torch.zeros(1, dtype=torch.int64, device="cuda")
    ck.set_traversal_stats(counter)
    try:
        ck.knn(torch.from_numpy(tree).cuda(), lay, torch.from_numpy(q).cuda(), 8, flags=flags)
        torch.cuda.synchronize()
    finally:
        ck.set_traversal_stats(None)
    assert int(counter.item()) == int(vis.sum())


@pytest.mark.parametrize("flags", [TWO, TWO | ck.QUERY_SORT])
def test_forest_two_per_lane(flags):
    """Per-lane queries on different trees, empty trees and empty segments
    included: identical to one query per lane."""
    D = 3
    sizes = log_uniform_sizes(80, 1, 30000, 44, empty_every=7)
    off = offsets_of(sizes)
    lay = ck.float_points(D)
    pts = torch.from_numpy(forest_points(sizes, D, 45)).cuda()
    f = ck.Forest(off, lay)
    f.build(pts)
    nq_t = (sizes // 3).astype(np.int64)
    nq_t[5::9] = 0
    nq_t[::13] += 3  # queries on empty trees too
    seg_off = torch.from_numpy(np.concatenate([[0], np.cumsum(nq_t)]).astype(np.int32)).cuda()
    seg_tree = torch.arange(len(sizes), dtype=torch.int32, device="cuda")
    q = torch.from_numpy(forest_points(np.maximum(nq_t, 0), D, 46)).cuda()
    for k in (1, 8):
        base = f.knn(pts, q, k, seg_off, seg_tree, flags=ONE)
        got = f.knn(pts, q, k, seg_off, seg_tree, flags=flags)
        torch.cuda.synchronize()
        for a, b in zip(got, base):
            assert torch.equal(a, b), k
    bi, bd = f.fcp(pts, q, seg_off, seg_tree, flags=ONE)
    gi, gd = f.fcp(pts, q, seg_off, seg_tree, flags=flags)
    torch.cuda.synchronize()
    assert torch.equal(bi, gi) and torch.equal(bd, gd)
    f.close()


def test_bench_shape_two_per_lane_sorted():
    """1M-point tree, bench query distribution, Morton order: the launch
    the benchmark measures."""
    s = _setup(datagen.uniform(1_000_000, 3, 1))
    q = datagen.uniform(20000, 3, 2)
    check_knn(*s, q, 8, float("inf"), lists=(ck.LIST_FIXED,), methods=SB,
              flags=TWO | ck.QUERY_SORT)


def test_large_sorted_and_presorted_launches_agree():
    """Batches of 2^19 queries and more launch the two-query kernel with
    fewer resident waves when sorted or marked presorted: results equal the
    plain launch (and the oracle on a subset)."""
    s = _setup(distribution("uniform", 200_000, 3, seed=51))
    lay, tree, box, dt, db = s
    q = datagen.uniform(600_000, 3, 52)
    dq = torch.from_numpy(q).cuda()
    order = ck.query_order(dq).long()
    qs = dq[order].contiguous()
    for k in (1, 4, 8, 16):
        plain = ck.knn(dt, lay, qs, k)
        pres = ck.knn(dt, lay, qs, k, flags=ck.QUERY_PRESORTED)
        srt = ck.knn(dt, lay, dq, k, flags=ck.QUERY_SORT)
        torch.cuda.synchronize()
        for a, b, c in zip(plain, pres, srt):
            assert torch.equal(a, b)
            assert torch.equal(a, c[order])
    fp = ck.fcp(dt, lay, qs)
    fs = ck.fcp(dt, lay, qs, flags=ck.QUERY_PRESORTED)
    torch.cuda.synchronize()
    assert torch.equal(fp[0], fs[0]) and torch.equal(fp[1], fs[1])
    sub = qs[:20000].cpu().numpy()
    check_knn(*s, sub, 8, float("inf"), lists=(ck.LIST_FIXED,), methods=SB,
              flags=ck.QUERY_PRESORTED)


@pytest.mark.parametrize("dist", ["coarse", "ties"])
def test_large_sorted_launch_with_ties(dist):
    """Tie-heavy points and queries at a large presorted batch (the launch
    that keeps pending plane distances in LDS): equal to the one-query-per-
    lane kernel, which the oracle pins on tied input elsewhere, and to the
    oracle on a presorted slice."""
    s = _setup(distribution(dist, 300_000, 3, seed=57))
    lay, tree, box, dt, db = s
    q = distribution(dist, 1 << 19, 3, seed=58) + np.float32(0.25)
    dq = torch.from_numpy(np.ascontiguousarray(q, dtype=np.float32)).cuda()
    qs = dq[ck.query_order(dq).long()].contiguous()
    for k in (4, 8):
        pres = ck.knn(dt, lay, qs, k, flags=ck.QUERY_PRESORTED)
        one = ck.knn(dt, lay, qs, k, flags=ONE)
        torch.cuda.synchronize()
        for a, b in zip(pres, one):
            assert torch.equal(a, b), k
    check_knn(*s, qs[:5000].cpu().numpy(), 8, float("inf"), lists=(ck.LIST_FIXED,),
              methods=SB, flags=ck.QUERY_PRESORTED)


def test_large_sorted_launch_statistics_exact():
    """The large presorted launch counts exactly the visits of the
    one-query-per-lane traversal (the reference's CUKD_STATS count)."""
    s = _setup(distribution("uniform", 400_000, 3, seed=61))
    lay, tree, box, dt, db = s
    dq = torch.from_numpy(datagen.uniform(1 << 19, 3, 62)).cuda()
    qs = dq[ck.query_order(dq).long()].contiguous()
    counts = []
    for flags in (ck.QUERY_PRESORTED, ONE):
        counter = torch.zeros(1, dtype=torch.int64, device="cuda")
        ck.set_traversal_stats(counter)
        try:
            ck.knn(dt, lay, qs, 8, flags=flags)
            torch.cuda.synchronize()
        finally:
            ck.set_traversal_stats(None)
        counts.append(int(counter.item()))
    assert counts[0] == counts[1] > 0
