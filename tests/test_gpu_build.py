"""GPU builder parity: the tree cukd_build_tree writes is byte-identical to
the oracle's buildTree_host restatement (builder_thrust.h:514-601), for every
builder entry point, element layout and tie pattern."""
import numpy as np
import pytest
import torch

import cudakdtree_amd as ck
import oracle as oc
from cudakdtree_amd import datagen
from tests.helpers import (check_kdtree_valid, coords_of, dims_of, distribution,
                           layout_for, make_elements)

pytestmark = pytest.mark.gpu

BUILDERS = [ck.BUILDER_DEFAULT, ck.BUILDER_THRUST, ck.BUILDER_BITONIC, ck.BUILDER_INPLACE]
# builders whose result equals buildTree_host on inputs with ties too (the
# in-place builder's tie rule is covered in test_gpu_builders_tl.py)
EXACT_BUILDERS = [ck.BUILDER_DEFAULT, ck.BUILDER_THRUST, ck.BUILDER_BITONIC]


def gpu_build(elems, layout, builder=ck.BUILDER_DEFAULT, want_bounds=False):
    t = torch.from_numpy(np.ascontiguousarray(elems)).cuda()
    b = ck.bounds_tensor(layout) if (want_bounds or layout.explicit_dim) else None
    ck.build_tree(t, layout, builder, b)
    torch.cuda.synchronize()
    return t.cpu().numpy(), (None if b is None else b.cpu().numpy())


def compare(coords, kind, builder=ck.BUILDER_DEFAULT, want_bounds=True):
    lay = layout_for(kind, coords.shape[1])
    elems = make_elements(coords, lay)
    ref, rbox = oc.build_host(elems, lay, want_bounds=True)
    got, gbox = gpu_build(elems, lay, builder, want_bounds)
    if not np.array_equal(got, ref):
        bad = np.nonzero(np.any(got != ref, axis=1))[0]
        raise AssertionError("%d of %d nodes differ, first at %d" % (len(bad), len(ref), bad[0]))
    if gbox is not None:
        # compare as scalars: -0.0 == +0.0 (box_t::grow keeps either)
        assert np.array_equal(gbox.view(rbox.dtype), rbox)
    return lay, got


@pytest.mark.parametrize("n", [1, 2, 3, 4, 7, 8, 9, 100, 511, 512, 513, 1000, 4095, 5000,
                               8191, 8192, 8193, 20000, 65537])
def test_float3_uniform_sizes(n):
    compare(datagen.uniform(n, 3, n), "float")


@pytest.mark.parametrize("builder", EXACT_BUILDERS)
@pytest.mark.parametrize("dist", ["uniform", "ties", "dups", "plane", "gmm", "signed_zero",
                                  "unique", "coarse"])
def test_float3_distributions(builder, dist):
    compare(distribution(dist, 30011, 3, seed=2), "float", builder)


@pytest.mark.parametrize("D", [1, 2, 3, 4, 5, 8])
@pytest.mark.parametrize("kind", ["float", "int", "payload", "payload_int"])
def test_layouts_round_robin(D, kind):
    scalar_float = kind in ("float", "payload")
    coords = distribution("ties" if D > 2 else "uniform", 12345, D, seed=D,
                          scalar_float=scalar_float)
    compare(coords, kind)


@pytest.mark.parametrize("D", [2, 3, 4, 8])
@pytest.mark.parametrize("kind", ["explicit", "explicit_int", "photon"])
@pytest.mark.parametrize("dist", ["uniform", "ties", "gmm"])
@pytest.mark.parametrize("n", [1, 6, 3000, 40000])
def test_explicit_dims(D, kind, dist, n):
    scalar_float = kind != "explicit_int"
    coords = distribution(dist, n, D, seed=n + D, scalar_float=scalar_float)
    lay, got = compare(coords, kind)
    check_kdtree_valid(coords_of(got, lay), dims_of(got, lay), n)


@pytest.mark.parametrize("D", [2, 3, 4, 8])
@pytest.mark.parametrize("kind", ["float", "int", "payload", "explicit", "explicit_int"])
@pytest.mark.parametrize("n", [4, 10, 1000, 10000])
def test_builders_same_result_matrix(D, kind, n):
    """testing/testBuildersSameResult.cu:404-430: tie-free unique-int
    coordinates, all builders equal the host builder."""
    c = datagen.unique_ints(n, D, seed=n + 7 * D)
    coords = c.astype(np.float32) if kind in ("float", "payload", "explicit") else c
    for b in BUILDERS:
        compare(coords, kind, b)


def test_two_million_uniform_and_gmm():
    for dist in ("uniform", "gmm", "coarse"):
        compare(distribution(dist, 2_000_000, 3, seed=77), "float")


@pytest.mark.parametrize("D", [1, 2, 4])
@pytest.mark.parametrize("kind", ["float", "int", "payload"])
def test_short_tie_runs(D, kind):
    """Short runs of equal coordinates inside the LDS-finished subtrees are
    ordered by the rest of the key (the bitonic and rank finishes must
    agree with the host builder)."""
    scalar_float = kind != "int"
    compare(distribution("coarse", 300_000, D, seed=5 + D, scalar_float=scalar_float), kind)


def test_repeat_builds_identical():
    coords = distribution("ties", 200_000, 3, seed=4)
    lay = ck.float_points(3)
    a, _ = gpu_build(coords, lay)
    b, _ = gpu_build(coords, lay)
    assert np.array_equal(a, b)


def test_user_stream_and_memory_resource():
    coords = datagen.uniform(100_000, 3, 5)
    lay = ck.float_points(3)
    ref, _ = oc.build_host(coords, lay)
    calls = {"malloc": 0, "free": 0}
    keep = {}

    @ck.MemResource._fields_[1][1]
    def _malloc(ctx, pptr, nbytes, stream):
        t = torch.empty(int(nbytes), dtype=torch.uint8, device="cuda")
        keep[t.data_ptr()] = t
        pptr[0] = t.data_ptr()
        calls["malloc"] += 1
        return 0

    @ck.MemResource._fields_[2][1]
    def _free(ctx, ptr, stream):
        calls["free"] += 1
        return 0

    mr = ck.MemResource(None, _malloc, _free)
    s = torch.cuda.Stream()
    t = torch.from_numpy(coords.copy()).cuda()
    torch.cuda.synchronize()
    import ctypes
    rc = ck.lib().cukd_build_tree(ctypes.c_void_p(t.data_ptr()), len(coords), ctypes.byref(lay),
                                  0, None, ctypes.c_void_p(s.cuda_stream), ctypes.byref(mr))
    assert rc == 0
    s.synchronize()
    assert calls["malloc"] == 1 and calls["free"] == 1
    assert np.array_equal(t.cpu().numpy(), ref)


def test_memory_resource_failure_reports_error():
    import ctypes
    lay = ck.float_points(3)

    @ck.MemResource._fields_[1][1]
    def _malloc(ctx, pptr, nbytes, stream):
        return 1

    @ck.MemResource._fields_[2][1]
    def _free(ctx, ptr, stream):
        return 0

    mr = ck.MemResource(None, _malloc, _free)
    t = torch.zeros((1000, 3), dtype=torch.float32, device="cuda")
    rc = ck.lib().cukd_build_tree(ctypes.c_void_p(t.data_ptr()), 1000, ctypes.byref(lay), 0,
                                  None, None, ctypes.byref(mr))
    assert rc == ck.ERR_ALLOC
    assert b"allocate" in ck.lib().cukd_last_error()


def test_compute_bounds():
    coords = distribution("gmm", 123457, 4, seed=3)
    lay = ck.float_points(4)
    b = ck.bounds_tensor(lay)
    ck.computeBounds(b, torch.from_numpy(coords).cuda(), lay)
    np.testing.assert_array_equal(b.cpu().numpy(), oc.compute_bounds(coords, lay))


def test_ten_million_builds_repeat_and_agree():
    """At the bench size (10M float3, 24-bit uniform coordinates, so ties
    are common) the radix-select builder gives the same array on every run,
    and that array equals the independent bitonic builder's; both builders
    equal buildTree_host at every size the oracle checks quickly.  Many
    repeats: a lost histogram count in the cross-workgroup select shows up
    in only a few builds out of a hundred."""
    n = 10_000_000
    coords = datagen.uniform(n, 3, 1)
    lay = ck.float_points(3)
    src = torch.from_numpy(coords).cuda()
    ref = src.clone()
    ck.build_tree(ref, lay, ck.BUILDER_BITONIC)
    t = torch.empty_like(src)
    for i in range(24):
        t.copy_(src)
        ck.build_tree(t, lay, ck.BUILDER_DEFAULT)
        assert torch.equal(t, ref), "default build %d differs from the bitonic build" % i


def _is_permutation(a, b):
    a = np.ascontiguousarray(a).view(np.uint32).reshape(len(a), -1)
    b = np.ascontiguousarray(b).view(np.uint32).reshape(len(b), -1)
    return np.array_equal(a[np.lexsort(a.T)], b[np.lexsort(b.T)])


@pytest.mark.parametrize("dist", ["uniform", "gmm"])
def test_ten_million_default_equals_oracle(dist):
    """The bench's tree size (10M float3; uniform 24-bit coordinates have
    millions of exactly tied values per axis, the 64-centre GMM is the
    clustered case): the default builder's array is byte-identical to the
    oracle's buildTree_host restatement (builder_thrust.h:514-601)."""
    n = 10_000_000
    coords = datagen.uniform(n, 3, 1) if dist == "uniform" else datagen.gmm(n, 3, 11)
    lay = ck.float_points(3)
    print("10M %s: GPU build" % dist, flush=True)
    got, gbox = gpu_build(coords, lay, want_bounds=True)
    print("10M %s: oracle build" % dist, flush=True)
    ref, rbox = oc.build_host(coords, lay, want_bounds=True)
    if not np.array_equal(got, ref):
        bad = np.nonzero(np.any(got != ref, axis=1))[0]
        raise AssertionError("%d of %d nodes differ, first at %d" % (len(bad), n, bad[0]))
    assert np.array_equal(gbox.view(rbox.dtype), rbox)


def test_ten_million_inplace_valid_permutation():
    """The in-place builder at 10M (ties included): a valid k-d tree over a
    permutation of the input, and the same array on a second run."""
    n = 10_000_000
    coords = datagen.uniform(n, 3, 1)
    lay = ck.float_points(3)
    print("10M in-place builds", flush=True)
    got, _ = gpu_build(coords, lay, ck.BUILDER_INPLACE)
    again, _ = gpu_build(coords, lay, ck.BUILDER_INPLACE)
    assert np.array_equal(got, again)
    lv = np.floor(np.log2(np.arange(n, dtype=np.float64) + 1)).astype(np.int64) % 3
    check_kdtree_valid(got, lv, n)
    assert _is_permutation(coords, got)


@pytest.mark.parametrize("kind", ["explicit", "photon"])
@pytest.mark.parametrize("n", [150_000, 300_000, 1_000_000])
def test_explicit_dims_large(kind, n):
    """Explicit-dim trees large enough that the finish runs on subtrees of
    1024..4096 elements (one tree: fewer than 512 finish workgroups), equal
    to the oracle."""
    coords = distribution("uniform" if n != 300_000 else "gmm", n, 3, seed=n)
    lay, got = compare(coords, kind)
    check_kdtree_valid(coords_of(got, lay), dims_of(got, lay), n)


@pytest.mark.parametrize("D", [2, 3, 4])
@pytest.mark.parametrize("kind", ["explicit", "explicit_int", "photon"])
@pytest.mark.parametrize("dist", ["grid17", "coarse"])
def test_explicit_dims_short_tie_runs(D, kind, dist):
    """Explicit-dim trees over a 2^-17 grid (a few percent of the
    coordinates in short runs of equal values: the rank finish repairs
    them level by level by the parent's split dim) and a 2^-14 grid (dense
    ties, and int coordinates with runs beyond the run limit: subtrees
    handed to the bitonic finish)."""
    coords = distribution(dist, 300_000, D, seed=31 + D, scalar_float=kind != "explicit_int")
    lay, got = compare(coords, kind)
    check_kdtree_valid(coords_of(got, lay), dims_of(got, lay), len(coords))


def test_explicit_dims_ten_million():
    """PointPlusPayload {float3; int payload; int dim} (20 bytes) at 10M:
    the default builder equals the bitonic builder (which is exact against
    the oracle at every size tested), and the tree is valid with every
    payload once."""
    n = 10_000_000
    coords = datagen.uniform(n, 3, 21)
    lay = ck.point_plus_payload(3, explicit_dim=True)
    elems = make_elements(coords, lay)
    elems[:, 3] = np.arange(n, dtype=np.uint32)
    print("10M explicit-dim builds", flush=True)
    a, abox = gpu_build(elems, lay, ck.BUILDER_DEFAULT)
    b, bbox = gpu_build(elems, lay, ck.BUILDER_BITONIC)
    assert np.array_equal(a, b)
    assert np.array_equal(abox, bbox)
    check_kdtree_valid(coords_of(a, lay), dims_of(a, lay), n)
    assert np.array_equal(np.sort(a[:, 3]), np.arange(n, dtype=np.uint32))


@pytest.mark.parametrize("builder", BUILDERS)
@pytest.mark.parametrize("kind", ["float", "explicit"])
def test_nan_rejected_array_unchanged(builder, kind):
    """A NaN coordinate has no place in the reference's '<' order
    (builder_thrust.h:401-404): every builder refuses it with
    CUKD_ERR_INVALID and leaves the array as it was."""
    coords = datagen.uniform(50_000, 3, 9)
    coords[31_337, 1] = np.nan
    lay = layout_for(kind, 3)
    elems = make_elements(coords, lay)
    t = torch.from_numpy(elems.copy()).cuda()
    b = ck.bounds_tensor(lay)
    with pytest.raises(ck.CukdError) as ei:
        ck.build_tree(t, lay, builder, b)
    assert ei.value.code == ck.ERR_INVALID and "NaN" in str(ei.value)
    torch.cuda.synchronize()
    assert np.array_equal(t.cpu().numpy(), elems)
    with pytest.raises(ck.CukdError):
        ck.buildTree_host(elems.copy(), lay, want_bounds=True)


@pytest.mark.parametrize("kind", ["float", "int", "explicit", "photon"])
def test_build_tree_host_equals_oracle(kind):
    """buildTree_host (builder_thrust.h:114-118) over host memory: the
    array and the host bounds box equal the oracle's."""
    coords = distribution("ties", 77_777, 3, seed=12, scalar_float=kind != "int")
    lay = layout_for(kind, 3)
    elems = make_elements(coords, lay)
    ref, rbox = oc.build_host(elems, lay, want_bounds=True)
    got = elems.copy()
    box = ck.buildTree_host(got, lay, want_bounds=True)
    assert np.array_equal(got, ref)
    assert np.array_equal(box.view(np.uint32), rbox.view(np.uint32))
    if kind == "explicit":
        with pytest.raises(ck.CukdError) as ei:
            ck.buildTree_host(elems.copy(), lay)
        assert ei.value.code == ck.ERR_NO_BOUNDS
