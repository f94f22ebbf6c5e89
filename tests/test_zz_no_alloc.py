"""cukd_knn_noalloc: the allocation-free kNN entry point (no counterpart in
the reference; for callers that cannot allocate).  Through the C ABI, its
results equal cukd_knn's bit for bit, on two calls in a row, the second
with fewer queries, Morton-ordered (the static-scratch path) and not."""
import ctypes

import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
from cudakdtree_amd import datagen

pytestmark = pytest.mark.gpu


def _call(fn, tree, lay, q, k, flags, list_kind=ck.LIST_FIXED, with_mem=True):
    nq = q.shape[0]
    ids = torch.full((nq, k), -7, dtype=torch.int32, device="cuda")
    d2 = torch.full((nq, k), -7.0, dtype=torch.float32, device="cuda")
    r = torch.full((nq,), -7.0, dtype=torch.float32, device="cuda")
    args = [ck.STACK_BASED, list_kind, k, ctypes.c_void_p(tree.data_ptr()), tree.shape[0],
            ctypes.byref(lay), None, ctypes.c_void_p(q.data_ptr()), nq, float("inf"),
            ctypes.c_void_p(ids.data_ptr()), ctypes.c_void_p(d2.data_ptr()),
            ctypes.c_void_p(r.data_ptr()), flags,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)]
    if with_mem:
        args.append(None)
    rc = fn(*args)
    assert rc == 0, ck.lib().cukd_last_error()
    torch.cuda.synchronize()
    return ids.cpu().numpy(), d2.cpu().numpy(), r.cpu().numpy()


@pytest.mark.parametrize("flags", [ck.QUERY_SORT, 0])
@pytest.mark.parametrize("k,list_kind", [(8, ck.LIST_FIXED), (20, ck.LIST_HEAP)])
def test_noalloc_equals_knn_two_calls(flags, k, list_kind):
    L = ck.lib()
    lay = ck.float_points(3)
    tree = torch.from_numpy(datagen.uniform(200_000, 3, 3)).cuda()
    ck.buildTree(tree, lay)
    for nq, seed in ((300_000, 5), (40_000, 6)):  # the second call smaller
        print("noalloc flags=%d k=%d nq=%d" % (flags, k, nq), flush=True)
        q = torch.from_numpy(datagen.uniform(nq, 3, seed)).cuda()
        want = _call(L.cukd_knn, tree, lay, q, k, flags, list_kind)
        got = _call(L.cukd_knn_noalloc, tree, lay, q, k, flags, list_kind, with_mem=False)
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1].view(np.uint32), want[1].view(np.uint32))
        assert np.array_equal(got[2].view(np.uint32), want[2].view(np.uint32))


def test_noalloc_rejects_more_queries_than_its_scratch():
    L = ck.lib()
    lay = ck.float_points(3)
    tree = torch.zeros((16, 3), dtype=torch.float32, device="cuda")
    q = torch.zeros((1, 3), dtype=torch.float32, device="cuda")
    out = torch.zeros(1, dtype=torch.int32, device="cuda")
    rc = L.cukd_knn_noalloc(ck.STACK_BASED, ck.LIST_FIXED, 4, ctypes.c_void_p(tree.data_ptr()),
                            16, ctypes.byref(lay), None, ctypes.c_void_p(q.data_ptr()),
                            (1 << 22) + 1, float("inf"), ctypes.c_void_p(out.data_ptr()),
                            ctypes.c_void_p(out.data_ptr()), None, ck.QUERY_SORT, None)
    assert rc == ck.ERR_INVALID
    assert b"static scratch" in L.cukd_last_error()
