"""The bitonic builder (n ints of temporary memory) and the in-place builder
(none), builder_bitonic.h:346-379 and builder_inplace.h:444-480.

buildTree_bitonic must equal the oracle's buildTree_host for every input.
buildTree_inPlace must equal it wherever the split coordinates decide, as on
the unique-int inputs of the reference's cross-builder test
(testing/testBuildersSameResult.cu:404-430); on inputs with ties it must be
a valid left-balanced k-d tree over the same elements, identical on every
run."""
import ctypes

import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.helpers import (check_kdtree_valid, coords_of, dims_of, distribution,
                           layout_for, make_elements, same_multiset)

pytestmark = pytest.mark.gpu

ALL_KINDS = ["float", "int", "payload", "payload_int", "explicit", "explicit_int", "photon",
             "rec96", "rec256"]


def run(elems, lay, builder):
    t = torch.from_numpy(np.ascontiguousarray(elems)).cuda()
    b = ck.bounds_tensor(lay)
    ck.build_tree(t, lay, builder, b)
    torch.cuda.synchronize()
    return t.cpu().numpy(), b.cpu().numpy()


def check(coords, kind, builder, exact):
    lay = layout_for(kind, coords.shape[1])
    elems = make_elements(coords, lay)
    ref, rbox = oc.build_host(elems, lay, want_bounds=True)
    got, gbox = run(elems, lay, builder)
    assert np.array_equal(gbox.view(rbox.dtype), rbox)
    if exact:
        if not np.array_equal(got, ref):
            bad = np.nonzero(np.any(got != ref, axis=1))[0]
            raise AssertionError("%d of %d nodes differ, first at %d" % (len(bad), len(ref), bad[0]))
    else:
        assert same_multiset(got, elems, lay)
        n = len(coords)
        if lay.explicit_dim:
            dims = dims_of(got, lay)
        else:
            dims = np.floor(np.log2(np.arange(n, dtype=np.float64) + 1)).astype(np.int64) % lay.num_dims
        check_kdtree_valid(coords_of(got, lay), dims, n)
        again, _ = run(elems, lay, builder)
        assert np.array_equal(again, got)
    return got


def unique_coords(n, D, kind, seed):
    c = datagen.unique_ints(n, D, seed=seed)
    return c if kind in ("int", "payload_int", "explicit_int") else c.astype(np.float32)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 100, 511, 512, 4095, 4096, 4097, 8191, 8192, 8193,
                               50000, 65537, 300001])
def test_sizes(n):
    check(distribution("ties", n, 3, seed=n), "float", ck.BUILDER_BITONIC, exact=True)
    check(unique_coords(n, 3, "float", n), "float", ck.BUILDER_INPLACE, exact=True)


@pytest.mark.parametrize("dist", ["uniform", "ties", "dups", "plane", "gmm", "signed_zero",
                                  "unique"])
def test_distributions(dist):
    c = distribution(dist, 30011, 3, seed=2)
    check(c, "float", ck.BUILDER_BITONIC, exact=True)
    check(c, "float", ck.BUILDER_INPLACE, exact=(dist == "unique"))


@pytest.mark.parametrize("D", [1, 2, 3, 4, 8])
@pytest.mark.parametrize("kind", ALL_KINDS)
@pytest.mark.parametrize("n", [3000, 40000])
def test_layouts(D, kind, n):
    scalar_float = kind not in ("int", "payload_int", "explicit_int")
    ties = distribution("ties" if D > 1 else "gmm", n, D, seed=n + D, scalar_float=scalar_float)
    check(ties, kind, ck.BUILDER_BITONIC, exact=True)
    check(unique_coords(n, D, kind, n * D), kind, ck.BUILDER_INPLACE, exact=True)
    if D in (2, 3):
        check(ties, kind, ck.BUILDER_INPLACE, exact=False)


def test_two_million():
    check(datagen.uniform(2_000_000, 3, 77), "float", ck.BUILDER_BITONIC, exact=True)
    check(unique_coords(2_000_000, 3, "float", 78), "float", ck.BUILDER_INPLACE, exact=True)


def _counting_resource():
    calls = {"malloc": 0, "free": 0, "bytes": 0}
    keep = {}

    @ck.MemResource._fields_[1][1]
    def _malloc(ctx, pptr, nbytes, stream):
        t = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        keep[t.data_ptr()] = t
        pptr[0] = t.data_ptr()
        calls["malloc"] += 1
        calls["bytes"] += int(nbytes)
        return 0

    @ck.MemResource._fields_[2][1]
    def _free(ctx, ptr, stream):
        calls["free"] += 1
        return 0

    return ck.MemResource(None, _malloc, _free), calls, (_malloc, _free, keep)


@pytest.mark.parametrize("builder", [ck.BUILDER_BITONIC, ck.BUILDER_INPLACE])
def test_temporary_memory(builder):
    """bitonic asks for exactly the n ints cukd_build_tree_temp_bytes
    reports; in-place asks for nothing (builder.h:34-45)."""
    n = 100_000
    lay = ck.float_points(3)
    coords = unique_coords(n, 3, "float", 5)
    ref, _ = oc.build_host(coords, lay)
    mr, calls, _hold = _counting_resource()
    t = torch.from_numpy(coords.copy()).cuda()
    torch.cuda.synchronize()
    rc = ck.lib().cukd_build_tree(ctypes.c_void_p(t.data_ptr()), n, ctypes.byref(lay), builder,
                                  None, None, ctypes.byref(mr))
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), ref)
    want = ck.build_temp_bytes(n, lay, builder)
    if builder == ck.BUILDER_INPLACE:
        assert want == 0 and calls["malloc"] == 0 and calls["free"] == 0
    else:
        assert want == 4 * n and calls["bytes"] == want and calls["free"] == calls["malloc"] == 1
