"""GPU spatial k-d tree parity (spatial-kdtree.h, fcp.h:256-427,
knn.h:566-752): nodes, primIDs and bounds equal the oracle's deterministic
build; queries equal the oracle's spatial traversals bit for bit."""
import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.helpers import distribution

pytestmark = pytest.mark.gpu


def build_both(coords, leaf=0):
    lay = ck.float_points(coords.shape[1])
    rn, rp, rb = oc.spatial_build(coords, lay, leaf)
    dp = torch.from_numpy(np.ascontiguousarray(coords)).cuda()
    t = ck.buildSpatialTree(dp, lay, leaf)
    return lay, (rn, rp, rb), t


@pytest.mark.parametrize("n", [0, 1, 8, 9, 100, 5000, 100000])
@pytest.mark.parametrize("dist", ["uniform", "gmm", "ties", "dups"])
def test_spatial_build_parity(n, dist):
    coords = distribution(dist, n, 3, seed=n + 1)
    lay, (rn, rp, rb), t = build_both(coords)
    assert t.num_nodes == len(rn)
    gn = t.nodes_numpy()
    assert np.array_equal(gn.view(np.uint8), rn.view(np.uint8))
    assert np.array_equal(t.prims_numpy(), rp)
    assert np.array_equal(np.array(t.bounds, np.float32), rb)
    ck.free(t)


@pytest.mark.parametrize("D", [2, 3, 4])
@pytest.mark.parametrize("leaf", [0, 1, 16])
def test_spatial_dims_and_leaf_threshold(D, leaf):
    coords = distribution("gmm", 20000, D, seed=D)
    lay, (rn, rp, rb), t = build_both(coords, leaf)
    assert np.array_equal(t.nodes_numpy().view(np.uint8), rn.view(np.uint8))
    assert np.array_equal(t.prims_numpy(), rp)
    ck.free(t)


@pytest.mark.parametrize("method", [ck.STACK_BASED, ck.CCT])
@pytest.mark.parametrize("dist", ["uniform", "gmm", "dups"])
def test_spatial_queries(method, dist):
    D = 3
    coords = distribution(dist, 30000, D, seed=3)
    lay, (rn, rp, rb), t = build_both(coords)
    q = datagen.uniform(3000, D, 7)
    dq = torch.from_numpy(q).cuda()
    for cutoff in (float("inf"), 0.05):
        gid, gd2 = ck.spatial_fcp(t, dq, method, cutoff)
        rid, rd2 = oc.spatial_fcp(method, rn, rp, rb, coords, lay, q, cutoff)
        torch.cuda.synchronize()
        assert np.array_equal(gid.cpu().numpy(), rid)
        assert np.array_equal(gd2.cpu().numpy().view(np.uint32), rd2.view(np.uint32))
        for lk in (ck.LIST_FIXED, ck.LIST_HEAP):
            for k in (1, 4, 16, 20):
                gid, gd2, gr = ck.spatial_knn(t, dq, k, method, lk, cutoff)
                rid, rd2, rr = oc.spatial_knn(method, lk, k, rn, rp, rb, coords, lay, q, cutoff)
                torch.cuda.synchronize()
                assert np.array_equal(gid.cpu().numpy(), rid)
                assert np.array_equal(gd2.cpu().numpy().view(np.uint32), rd2.view(np.uint32))
                assert np.array_equal(gr.cpu().numpy().view(np.uint32), rr.view(np.uint32))
    ck.free(t)


def test_spatial_deep_stack_beyond_lds_entries():
    """The batched stack-based query keeps its first 16 stack entries in
    LDS and the rest in a per-lane array.  Unbounded-radius queries push a
    far child at every level of their first descent, and the leaves of this
    1M-point tree lie at depth 21 (median; 6 to 26), so both parts are
    used."""
    D = 3
    coords = distribution("gmm", 1_000_000, D, seed=21)
    lay, (rn, rp, rb), t = build_both(coords)
    q = datagen.uniform(1500, D, 22)
    dq = torch.from_numpy(q).cuda()
    cutoff = float("inf")
    gid, gd2 = ck.spatial_fcp(t, dq, ck.STACK_BASED, cutoff)
    rid, rd2 = oc.spatial_fcp(oc.STACK_BASED, rn, rp, rb, coords, lay, q, cutoff)
    torch.cuda.synchronize()
    assert np.array_equal(gid.cpu().numpy(), rid)
    assert np.array_equal(gd2.cpu().numpy().view(np.uint32), rd2.view(np.uint32))
    for lk, k in ((ck.LIST_FIXED, 4), (ck.LIST_FIXED, 16), (ck.LIST_HEAP, 16), (ck.LIST_FIXED, 33)):
        gid, gd2, gr = ck.spatial_knn(t, dq, k, ck.STACK_BASED, lk, cutoff)
        rid, rd2, rr = oc.spatial_knn(oc.STACK_BASED, lk, k, rn, rp, rb, coords, lay, q, cutoff)
        torch.cuda.synchronize()
        assert np.array_equal(gid.cpu().numpy(), rid), (lk, k)
        assert np.array_equal(gd2.cpu().numpy().view(np.uint32), rd2.view(np.uint32)), (lk, k)
        assert np.array_equal(gr.cpu().numpy().view(np.uint32), rr.view(np.uint32)), (lk, k)
    ck.free(t)


def test_spatial_knn16_cutoff_config_shape():
    """4M float4 in [-1,1)^4 with 1% duplicates; kNN k=16, r=0.05 on queries
    half from the data (+N(0,0.001) jitter) and half uniform -- subset of the
    queries for the oracle's time budget."""
    n, D = 4_000_000, 4
    coords = datagen.uniform(n, D, 11, -1.0, 1.0)
    dup = (datagen.splitmix64(12, n // 100) % np.uint64(n)).astype(np.int64)
    coords[n - n // 100:] = coords[dup]
    lay, (rn, rp, rb), t = build_both(coords)
    assert np.array_equal(t.prims_numpy(), rp)
    nq = 4000
    pick = (datagen.splitmix64(13, nq // 2) % np.uint64(n)).astype(np.int64)
    jit = (datagen.gaussian(nq // 2 * D, 14).reshape(nq // 2, D) * 0.001).astype(np.float32)
    q = np.concatenate([coords[pick] + jit, datagen.uniform(nq // 2, D, 15, -1.0, 1.0)]).astype(np.float32)
    dq = torch.from_numpy(q).cuda()
    gid, gd2, gr = ck.spatial_knn(t, dq, 16, ck.STACK_BASED, ck.LIST_HEAP, 0.05)
    rid, rd2, rr = oc.spatial_knn(oc.STACK_BASED, oc.LIST_HEAP, 16, rn, rp, rb, coords, lay, q, 0.05)
    torch.cuda.synchronize()
    assert np.array_equal(gid.cpu().numpy(), rid)
    assert np.array_equal(gd2.cpu().numpy().view(np.uint32), rd2.view(np.uint32))
    ck.free(t)


def _int3_coords(n, seed, dup_every=0):
    """int3 coordinates uniform in [0, 2^20) (the spatial config's int
    set), optionally with exact duplicates."""
    c = (datagen.splitmix64(seed, n * 3) >> np.uint64(44)).astype(np.int32).reshape(n, 3)
    if dup_every:
        m = n // dup_every
        src = (datagen.splitmix64(seed + 1, m) % np.uint64(n)).astype(np.int64)
        c[n - m:] = c[src]
    return c


@pytest.mark.parametrize("n", [1, 9, 1000, 100_000, 2_000_000])
def test_spatial_build_int3(n):
    """Spatial trees over int points (Node::pos is scalar_t,
    spatial-kdtree.h:69-98): nodes (int32 pos), primIDs and bounds equal
    the oracle's; 2M int3 with 1% duplicates is the config's int set."""
    coords = _int3_coords(n, 31 + n, dup_every=100 if n >= 1000 else 0)
    lay = ck.int_points(3)
    rn, rp, rb = oc.spatial_build(coords, lay)
    t = ck.buildSpatialTree(torch.from_numpy(coords).cuda(), lay)
    assert t.num_nodes == len(rn)
    assert np.array_equal(t.nodes_numpy().view(np.uint8), rn.view(np.uint8))
    assert np.array_equal(t.prims_numpy(), rp)
    assert np.array_equal(np.array(t.bounds, np.int32), rb)
    if n == 1000:
        with pytest.raises(ck.CukdError) as ei:
            ck.spatial_fcp(t, torch.zeros((4, 3), dtype=torch.float32, device="cuda"))
        assert ei.value.code == ck.ERR_UNSUPPORTED
    ck.free(t)
