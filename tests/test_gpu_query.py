"""GPU query parity: batched FCP / kNN over all three traversals and both
candidate lists return the oracle's ids, dist2 bits and return values
exactly (fcp.h, knn.h, traverse-*.h)."""
import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.helpers import distribution, layout_for, make_elements

pytestmark = pytest.mark.gpu

METHODS = [ck.STACK_BASED, ck.STACK_FREE, ck.CCT]


def _setup(coords, kind="float"):
    lay = layout_for(kind, coords.shape[1])
    elems = make_elements(coords, lay) if kind != "float" else coords
    tree, box = oc.build_host(elems, lay, want_bounds=True)
    dt = torch.from_numpy(np.ascontiguousarray(tree)).cuda()
    db = torch.from_numpy(box.astype(np.float32)).cuda()
    return lay, tree, box, dt, db


def _eq(a, b, what):
    if not np.array_equal(a, b):
        bad = np.nonzero(np.atleast_2d(a != b).any(axis=-1) if a.ndim > 1 else a != b)[0]
        raise AssertionError("%s differs at %d rows, first %d" % (what, len(bad), bad[0]))


def check_fcp(lay, tree, box, dt, db, q, cutoff, eps=0.0, methods=METHODS, flags=0):
    dq = torch.from_numpy(q).cuda()
    for m in methods:
        rid, rd2 = oc.fcp(m, tree, lay, q, cutoff=cutoff, eps=eps, world_bounds=box)
        gid, gd2 = ck.fcp(dt, lay, dq, m, cutoff, eps, db, flags=flags)
        torch.cuda.synchronize()
        _eq(gid.cpu().numpy(), rid, "fcp ids (method %d)" % m)
        _eq(gd2.cpu().numpy().view(np.uint32), rd2.view(np.uint32), "fcp dist2 (method %d)" % m)


def check_knn(lay, tree, box, dt, db, q, k, cutoff, lists=(ck.LIST_FIXED, ck.LIST_HEAP),
              methods=METHODS, flags=0):
    dq = torch.from_numpy(q).cuda()
    for m in methods:
        for lk in lists:
            rid, rd2, rret = oc.knn(m, lk, k, tree, lay, q, cutoff=cutoff, world_bounds=box)
            gid, gd2, gret = ck.knn(dt, lay, dq, k, m, lk, cutoff, db, flags=flags)
            torch.cuda.synchronize()
            tag = "(method %d list %d k %d)" % (m, lk, k)
            _eq(gid.cpu().numpy(), rid, "knn ids " + tag)
            _eq(gd2.cpu().numpy().view(np.uint32), rd2.view(np.uint32), "knn dist2 " + tag)
            _eq(gret.cpu().numpy().view(np.uint32), rret.view(np.uint32), "knn ret " + tag)


@pytest.mark.parametrize("D", [1, 2, 3, 4, 5, 8])
@pytest.mark.parametrize("dist", ["uniform", "gmm", "ties"])
def test_fcp_all_methods(D, dist):
    coords = distribution(dist, 20000, D, seed=D)
    s = _setup(coords)
    q = datagen.uniform(3000, D, 99) * (7.0 if dist == "ties" else 1.0)
    check_fcp(*s, q, float("inf"))
    check_fcp(*s, q, 0.02)


@pytest.mark.parametrize("k", [1, 2, 4, 5, 8, 16, 20, 50, 64])
def test_knn_k_values(k):
    coords = distribution("uniform", 30000, 3, seed=k)
    s = _setup(coords)
    q = datagen.uniform(2000, 3, k + 1)
    check_knn(*s, q, k, float("inf"))


@pytest.mark.parametrize("k", [20, 64])
def test_knn_run_time_k_lists_sorted_and_autotuned(k):
    """The run-time-k lists keep their entries in LDS columns, one column
    per lane, the workgroup size apart: Morton-ordered launches and the
    autotuner's 64/128/256-lane workgroups give the oracle's lists too."""
    coords = distribution("gmm", 40000, 3, seed=k + 5)
    s = _setup(coords)
    q = datagen.uniform(5000, 3, k + 6)
    check_knn(*s, q, k, float("inf"), flags=ck.QUERY_SORT)
    check_knn(*s, q, k, 0.1, flags=ck.QUERY_AUTOTUNE)


@pytest.mark.parametrize("D", [2, 3, 4, 8])
@pytest.mark.parametrize("dist", ["ties", "gmm", "dups"])
def test_knn_ties_and_cutoff(D, dist):
    coords = distribution(dist, 8000, D, seed=D + 1)
    s = _setup(coords)
    q = datagen.uniform(1500, D, 3) * (7.0 if dist == "ties" else 1.0)
    for cutoff in (float("inf"), 0.05, 1.0):
        check_knn(*s, q, 8, cutoff)


def test_stack_free_eps():
    coords = distribution("gmm", 20000, 3, seed=5)
    s = _setup(coords)
    q = datagen.uniform(3000, 3, 6)
    check_fcp(*s, q, float("inf"), eps=0.5, methods=[ck.STACK_FREE])


@pytest.mark.parametrize("kind", ["explicit", "photon"])
def test_explicit_dim_trees(kind):
    coords = distribution("gmm", 20000, 3, seed=8)
    s = _setup(coords, kind)
    q = datagen.uniform(3000, 3, 9)
    check_fcp(*s, q, float("inf"))
    check_knn(*s, q, 8, float("inf"))


def test_payload_layout():
    coords = distribution("uniform", 10000, 4, seed=8)
    s = _setup(coords, "payload")
    q = datagen.uniform(2000, 4, 9)
    check_fcp(*s, q, float("inf"))
    check_knn(*s, q, 4, float("inf"), lists=(ck.LIST_FIXED,))


def test_issue5_known_answer_gpu():
    pts = datagen.msvc_rand_points(90167, 33)
    lay = ck.float_points(3)
    dp = torch.from_numpy(pts).cuda()
    b = ck.bounds_tensor(lay)
    ck.buildTree(dp, lay, b)
    q = torch.tensor([[-98.4496613, 76.9219055, 25.8888512]], dtype=torch.float32, device="cuda")
    for m in METHODS:
        ids, d2 = ck.fcp(dp, lay, q, m, 5.0, 0.0, b)
        torch.cuda.synchronize()
        assert d2.cpu().numpy().view(np.uint32)[0] == 0x40362F30


def test_empty_tree_and_queries():
    lay = ck.float_points(3)
    dt = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    db = torch.tensor([0, 0, 0, 1, 1, 1], dtype=torch.float32, device="cuda")
    q = torch.from_numpy(datagen.uniform(10, 3, 1)).cuda()
    for m in METHODS:
        ids, d2 = ck.fcp(dt, lay, q, m, float("inf"), 0.0, db)
        torch.cuda.synchronize()
        assert np.all(ids.cpu().numpy() == -1)
    q0 = torch.zeros((0, 3), dtype=torch.float32, device="cuda")
    ck.fcp(dt, lay, q0)


def test_bench_shape_subset():
    """1M points, queries from the bench distribution (subset for the
    oracle's time budget)."""
    coords = datagen.uniform(1_000_000, 3, 1)
    s = _setup(coords)
    q = datagen.uniform(20000, 3, 2)
    check_knn(*s, q, 8, float("inf"), lists=(ck.LIST_FIXED,))
    check_fcp(*s, q, float("inf"))


@pytest.mark.parametrize("D", [2, 3, 4])
def test_morton_ordered_queries_identical(D):
    """CUKD_QUERY_SORT only changes the launch order: results equal the
    oracle's and land at each query's own index."""
    coords = distribution("gmm", 50000, D, seed=21)
    s = _setup(coords)
    q = datagen.uniform(20000, D, 22)
    check_fcp(*s, q, float("inf"), flags=ck.QUERY_SORT)
    check_knn(*s, q, 8, float("inf"), flags=ck.QUERY_SORT)
    check_knn(*s, q, 5, 0.05, lists=(ck.LIST_HEAP,), flags=ck.QUERY_SORT)
