"""Serving loop (cukd_server_*): requests batched by the worker thread give
exactly the results of the single-tree queries, whether a request is
served alone, in a burst with others, or with one launch per request."""
import threading

import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.test_gpu_forest import forest_points, log_uniform_sizes, offsets_of

pytestmark = pytest.mark.gpu

KINDS = [(0, np.inf), (1, np.inf), (4, np.inf), (16, np.inf), (4, 0.05), (0, 0.02)]


@pytest.fixture(scope="module")
def served_forest():
    D = 3
    sizes = log_uniform_sizes(40, 10, 200000, 51)
    off = offsets_of(sizes)
    lay = ck.float_points(D)
    pts = torch.from_numpy(forest_points(sizes, D, 53, dist="gmm")).cuda()
    f = ck.Forest(off, lay)
    f.build(pts)
    torch.cuda.synchronize()
    yield f, pts, off, lay
    f.close()


def _requests(off, count, seed):
    T = len(off) - 1
    r = datagen.splitmix64(seed, 4 * count)
    reqs = []
    for i in range(count):
        tree = int(r[4 * i] % np.uint64(T))
        u = float(r[4 * i + 1] >> np.uint64(11)) / float(1 << 53)
        nq = int(np.exp(u * np.log(20000)))  # 1 .. 20000, log-uniform
        k, cutoff = KINDS[int(r[4 * i + 2] % np.uint64(len(KINDS)))]
        q = datagen.gmm(nq, 3, seed * 1000 + i)
        reqs.append((tree, q, k, cutoff))
    return reqs


def _expected(pts, off, lay, tree, q, k, cutoff):
    t = pts[off[tree]:off[tree + 1]]
    dq = torch.from_numpy(q).cuda()
    if k == 0:
        ids, d2 = ck.fcp(t, lay, dq, cutoff=cutoff)
    else:
        ids, d2, _ = ck.knn(t, lay, dq, k, cutoff=cutoff)
    return ids.cpu().numpy(), d2.cpu().numpy()


def _serve(srv, reqs, one_at_a_time=False):
    tickets = []
    for tree, q, k, cutoff in reqs:
        tickets.append(srv.submit(tree, q, k, cutoff))
        if one_at_a_time:
            done = srv.poll(wait_for=1)
            assert len(done) == 1 and done[0].ticket == tickets[-1] and done[0].status == 0
    pending = set(tickets) if not one_at_a_time else set()
    while pending:
        for c in srv.poll(wait_for=1):
            assert c.status == 0
            pending.discard(c.ticket)
    return [srv.result(t) for t in tickets]


def test_burst_matches_single_tree_queries(served_forest):
    f, pts, off, lay = served_forest
    reqs = _requests(off, 60, 7)
    srv = ck.Server(f, pts, budget_bytes=64 << 20)
    got = _serve(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st.requests_done == len(reqs) and st.batches < len(reqs)
    for (tree, q, k, cutoff), (ids, d2) in zip(reqs, got):
        eids, ed2 = _expected(pts, off, lay, tree, q, k, cutoff)
        assert np.array_equal(ids, eids) and np.array_equal(d2.view(np.uint32), ed2.view(np.uint32))


def test_same_result_alone_and_in_the_mix(served_forest):
    f, pts, off, lay = served_forest
    reqs = _requests(off, 40, 9)
    srv = ck.Server(f, pts, budget_bytes=32 << 20)
    mixed = _serve(srv, reqs)
    alone = _serve(srv, reqs, one_at_a_time=True)
    srv.close()
    per_request = ck.Server(f, pts, budget_bytes=32 << 20, max_requests_per_batch=1)
    single = _serve(per_request, reqs)
    st = per_request.stats()
    per_request.close()
    assert st.batches == len(reqs)
    for a, b, c in zip(mixed, alone, single):
        for x, y, z in zip(a, b, c):
            assert np.array_equal(x, y) and np.array_equal(x, z)


def test_concurrent_submitters(served_forest):
    f, pts, off, lay = served_forest
    srv = ck.Server(f, pts, budget_bytes=64 << 20)
    all_reqs = [_requests(off, 15, 100 + t) for t in range(4)]
    results = [None] * 4
    lock = threading.Lock()

    def run(t):
        tickets = []
        for tree, q, k, cutoff in all_reqs[t]:
            with lock:
                tickets.append(srv.submit(tree, q, k, cutoff))
        results[t] = tickets

    th = [threading.Thread(target=run, args=(t,)) for t in range(4)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    want = sum(len(r) for r in all_reqs)
    seen = 0
    while seen < want:
        seen += len(srv.poll(wait_for=1))
    for t in range(4):
        for (tree, q, k, cutoff), ticket in zip(all_reqs[t], results[t]):
            ids, d2 = srv.result(ticket)
            eids, ed2 = _expected(pts, off, lay, tree, q, k, cutoff)
            assert np.array_equal(ids, eids) and np.array_equal(d2, ed2)
    srv.close()


def _oracle(pts_h, off, lay, tree, q, k, cutoff):
    t = pts_h[off[tree]:off[tree + 1]]
    if k == 0:
        return oc.fcp(oc.STACK_BASED, t, lay, q, cutoff=cutoff)
    ids, d2, _ = oc.knn(oc.STACK_BASED, oc.LIST_FIXED, k, t, lay, q, cutoff=cutoff)
    return ids, d2


def test_burst_matches_oracle(served_forest):
    """Every served request equals the oracle's stack-based fcp / knn
    (FixedCandidateList<k>) on that request's own tree: ids and dist2
    bits."""
    f, pts, off, lay = served_forest
    pts_h = pts.cpu().numpy()
    reqs = _requests(off, 40, 17)
    srv = ck.Server(f, pts, budget_bytes=64 << 20)
    got = _serve(srv, reqs)
    srv.close()
    for (tree, q, k, cutoff), (ids, d2) in zip(reqs, got):
        eids, ed2 = _oracle(pts_h, off, lay, tree, q, k, cutoff)
        assert np.array_equal(ids.ravel(), eids.ravel())
        assert np.array_equal(d2.view(np.uint32).ravel(), ed2.view(np.uint32).ravel())


def test_small_budget_many_classes(served_forest):
    """A 1 MiB budget (128 KiB slots) and hundreds of one-query requests,
    each with a cutoff of its own (one class each): admission charges the
    exact padded layout, so every batch fits and every request succeeds."""
    f, pts, off, lay = served_forest
    pts_h = pts.cpu().numpy()
    srv = ck.Server(f, pts, budget_bytes=1 << 20)
    T = len(off) - 1
    reqs = []
    for i in range(400):
        k = (0, 1, 4, 16)[i % 4]
        reqs.append((i % T, datagen.gmm(1, 3, 900 + i), k, 0.5 + i * 1e-3))
    got = _serve(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st.requests_done == len(reqs) and st.batches < len(reqs)
    for (tree, q, k, cutoff), (ids, d2) in zip(reqs, got):
        eids, ed2 = _oracle(pts_h, off, lay, tree, q, k, cutoff)
        assert np.array_equal(ids.ravel(), eids.ravel())
        assert np.array_equal(d2.view(np.uint32).ravel(), ed2.view(np.uint32).ravel())


def test_request_larger_than_a_slot_is_served_in_pieces(served_forest):
    """A request that does not fit one slot is queued in pieces, not
    refused, and its results equal the single-tree query."""
    f, pts, off, lay = served_forest
    srv = ck.Server(f, pts, budget_bytes=1 << 20)
    big = datagen.gmm(30000, 3, 5)          # 30000 * (12 + 128) bytes >> 128 KiB
    light = [(1, datagen.gmm(50, 3, 6), 4, np.inf), (2, datagen.gmm(7, 3, 8), 0, np.inf)]
    reqs = [(0, big, 16, np.inf)] + light
    got = _serve(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st.requests_done == 3 and st.batches >= 3
    for (tree, q, k, cutoff), (ids, d2) in zip(reqs, got):
        eids, ed2 = _expected(pts, off, lay, tree, q, k, cutoff)
        assert np.array_equal(ids.ravel(), eids.ravel())
        assert np.array_equal(d2.view(np.uint32).ravel(), ed2.view(np.uint32).ravel())


def test_light_and_heavy_requests_in_one_batch(served_forest):
    """Light requests and heavy ones (nq * k >= 65536) queued together: the
    light requests of a class share a launch and go first, then each heavy
    request has a launch of its own, cheapest first.  Every result equals
    the single-tree query."""
    f, pts, off, lay = served_forest
    srv = ck.Server(f, pts, budget_bytes=256 << 20)
    reqs = [(0, datagen.gmm(10, 3, 1), 4, np.inf), (3, datagen.gmm(20000, 3, 2), 4, np.inf),
            (1, datagen.gmm(10, 3, 3), 4, np.inf), (2, datagen.gmm(20, 3, 4), 0, np.inf),
            (4, datagen.gmm(9000, 3, 5), 16, np.inf), (5, datagen.gmm(30, 3, 6), 16, np.inf),
            (6, datagen.gmm(17000, 3, 7), 4, np.inf)]
    got = _serve(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st.requests_done == len(reqs)
    for (tree, q, k, cutoff), (ids, d2) in zip(reqs, got):
        eids, ed2 = _expected(pts, off, lay, tree, q, k, cutoff)
        assert np.array_equal(ids.ravel(), eids.ravel())
        assert np.array_equal(d2.view(np.uint32).ravel(), ed2.view(np.uint32).ravel())


def test_heavy_requests_beyond_one_launch_bound(served_forest):
    """Six heavy kNN-16 requests of one class submitted back to back (1.6M
    list slots each; six exceed the 2^23 slots one heavy launch takes):
    whatever the worker admits into one batch is packed into launches of
    bounded work, cheapest first, each completing on its own.  Every
    result equals the single-tree query."""
    f, pts, off, lay = served_forest
    srv = ck.Server(f, pts, budget_bytes=2 << 30)
    T = len(off) - 1
    reqs = [((3 * i + 1) % T, datagen.gmm(100_000 - 7 * i, 3, 300 + i), 16, np.inf)
            for i in range(6)]
    reqs.insert(2, (5 % T, datagen.gmm(40, 3, 299), 16, np.inf))  # a light one among them
    got = _serve(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st.requests_done == len(reqs)
    for (tree, q, k, cutoff), (ids, d2) in zip(reqs, got):
        eids, ed2 = _expected(pts, off, lay, tree, q, k, cutoff)
        assert np.array_equal(ids.ravel(), eids.ravel())
        assert np.array_equal(d2.view(np.uint32).ravel(), ed2.view(np.uint32).ravel())


def test_budget_and_argument_errors(served_forest):
    f, pts, off, lay = served_forest
    srv = ck.Server(f, pts, budget_bytes=1 << 20)
    with pytest.raises(ck.CukdError):
        srv.submit(len(off) - 1, datagen.uniform(10, 3, 1), 4)  # no such tree
    with pytest.raises(ck.CukdError):
        srv.submit(0, datagen.uniform(10, 3, 1), 65)
    # pending work is served before close returns
    tickets = [srv.submit(i % (len(off) - 1), datagen.uniform(500, 3, i), 4) for i in range(20)]
    srv.close()
    assert all(t in srv._keep for t in tickets)


def test_caller_result_buffers(served_forest):
    """submit(out=...) writes into the caller's arrays, with the same results
    as server-allocated ones; wrong shapes are refused."""
    f, pts, off, lay = served_forest
    reqs = _requests(off, 12, 11)
    srv = ck.Server(f, pts, budget_bytes=32 << 20)
    ref = _serve(srv, reqs)
    outs, tickets = [], []
    for tree, q, k, cutoff in reqs:
        n = len(q) * max(1, k)
        outs.append((np.full(n, -5, np.int32), np.full(n, -5.0, np.float32)))
        tickets.append(srv.submit(tree, q, k, cutoff, out=outs[-1]))
    pending = set(tickets)
    while pending:
        for c in srv.poll(wait_for=1):
            assert c.status == 0
            pending.discard(c.ticket)
    for t, (ids, d2), (rids, rd2) in zip(tickets, outs, ref):
        got_ids, got_d2 = srv.result(t)
        assert got_ids is ids and got_d2 is d2
        assert np.array_equal(ids, rids.ravel()) and np.array_equal(d2, rd2.ravel())
    with pytest.raises(ValueError):
        srv.submit(0, reqs[0][1], 4, out=(np.zeros(3, np.int32), np.zeros(3, np.float32)))
    srv.close()


def test_server_starts_after_the_creation_stream(served_forest):
    """The server's batch streams wait for the work queued on the stream
    given at creation: a forest built there, with no host synchronisation
    before the server is created, is served with the single-tree results."""
    f, pts, off, lay = served_forest
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        pts2 = pts.clone()
        f2 = ck.Forest(off, lay)
        f2.build(pts2, stream=side)
    srv = ck.Server(f2, pts2, budget_bytes=32 << 20, stream=side)
    reqs = _requests(off, 16, 13)
    got = _serve(srv, reqs)
    srv.close()
    torch.cuda.synchronize()
    for (tree, q, k, cutoff), (ids, d2) in zip(reqs, got):
        eids, ed2 = _expected(pts2, off, lay, tree, q, k, cutoff)
        assert np.array_equal(ids, eids) and np.array_equal(d2.view(np.uint32), ed2.view(np.uint32))
    f2.close()
