"""Shared test helpers: layouts, point-set construction, GPU transfer."""
import numpy as np

import cudakdtree_amd as ck
from cudakdtree_amd import datagen


def layout_for(kind, d):
    """kind: 'float', 'int', 'payload', 'payload_int', 'explicit',
    'explicit_int', 'photon' (float3 + uint8 dim at byte 12, stride 16),
    'rec96' (96-byte record, point at byte 32, no dim), 'rec256' (256-byte
    record, point at byte 16, int32 dim at byte 200)."""
    if kind == "float":
        return ck.float_points(d)
    if kind == "int":
        return ck.int_points(d)
    if kind == "payload":
        return ck.point_plus_payload(d)
    if kind == "payload_int":
        return ck.point_plus_payload(d, ck.SCALAR_I32)
    if kind == "explicit":
        return ck.point_plus_payload(d, explicit_dim=True)
    if kind == "explicit_int":
        return ck.point_plus_payload(d, ck.SCALAR_I32, explicit_dim=True)
    if kind == "photon":
        return ck.Layout(d, ck.SCALAR_F32, 4 * d + 4, 0, 4 * d, 1)
    if kind == "rec96":
        return ck.Layout(d, ck.SCALAR_F32, 96, 32, -1, 4)
    if kind == "rec256":
        return ck.Layout(d, ck.SCALAR_F32, 256, 16, 200, 4)
    raise ValueError(kind)


def make_elements(coords, layout, seed=7):
    """Packs coordinates (n, D) into uint32 words of layout.stride bytes;
    payload words get distinct values, the dim field starts as garbage."""
    n, D = coords.shape
    W = layout.stride // 4
    out = np.zeros((n, W), np.uint32)
    po = layout.point_offset // 4
    out[:, po:po + D] = np.ascontiguousarray(coords).view(np.uint32).reshape(n, D)
    for w in range(W):
        if po <= w < po + D:
            continue
        out[:, w] = (datagen.splitmix64(seed + w, n) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    if layout.dim_offset >= 0:
        raw = out.view(np.uint8).reshape(n, layout.stride)
        raw[:, layout.dim_offset:layout.dim_offset + layout.dim_bytes] = 0xAB
    return out


def coords_of(elems, layout):
    D = layout.num_dims
    po = layout.point_offset // 4
    w = np.ascontiguousarray(elems[:, po:po + D])
    return w.view(np.float32 if layout.scalar_type == ck.SCALAR_F32 else np.int32)


def dims_of(elems, layout):
    raw = np.ascontiguousarray(elems).view(np.uint8).reshape(len(elems), layout.stride)
    b = raw[:, layout.dim_offset:layout.dim_offset + layout.dim_bytes].astype(np.int64)
    v = np.zeros(len(elems), np.int64)
    for i in range(layout.dim_bytes):
        v |= b[:, i] << (8 * i)
    return v


def distribution(name, n, d, seed, scalar_float=True):
    if name == "uniform":
        c = datagen.uniform(n, d, seed)
    elif name == "ties":      # few distinct values per axis
        c = (datagen.splitmix64(seed, n * d) % np.uint64(7)).astype(np.float32).reshape(n, d)
    elif name == "dups":      # every point equal
        c = np.full((n, d), 0.5, np.float32)
    elif name == "plane":     # last axis constant, the rest uniform
        c = datagen.uniform(n, d, seed)
        c[:, -1] = 0.25
    elif name == "gmm":
        c = datagen.gmm(n, d, seed)
    elif name == "coarse":    # uniform on a 2^-14 grid: many short runs of equal values
        c = np.floor(datagen.uniform(n, d, seed) * 16384.0).astype(np.float32) / np.float32(16384.0)
    elif name == "grid17":    # uniform on a 2^-17 grid: a few percent of the values tied
        c = np.floor(datagen.uniform(n, d, seed) * 131072.0).astype(np.float32) / np.float32(131072.0)
    elif name == "unique":
        c = datagen.unique_ints(n, d, seed).astype(np.float32)
    elif name == "signed_zero":  # -0.0 and +0.0 must compare equal
        c = (datagen.splitmix64(seed, n * d) % np.uint64(3)).astype(np.float32).reshape(n, d) - 1.0
        neg = (datagen.splitmix64(seed ^ 99, n * d) & np.uint64(1)).astype(bool).reshape(n, d)
        c[(c == 0) & neg] = -0.0
    else:
        raise ValueError(name)
    if not scalar_float:
        c = np.round(c * 1000).astype(np.int32)
    return c


def same_multiset(a, b, layout):
    """a and b hold the same elements, ignoring the dim field."""
    def canon(x):
        x = np.ascontiguousarray(x).copy()
        if layout.dim_offset >= 0:
            raw = x.view(np.uint8).reshape(len(x), layout.stride)
            raw[:, layout.dim_offset:layout.dim_offset + layout.dim_bytes] = 0
        return x[np.lexsort(x.T[::-1])]
    return np.array_equal(canon(a), canon(b))


def check_kdtree_valid(coords, dims, n):
    """Every node lies on the correct side of each ancestor's split plane
    (testing/floatN-knn-and-fcp.cu:387-427, vectorised)."""
    i = np.arange(n, dtype=np.int64)
    j = 1
    while True:
        anc = ((i + 1) >> j) - 1
        sel = anc >= 0
        if not sel.any():
            break
        ii, aa = i[sel], anc[sel]
        right = (((ii + 1) >> (j - 1)) & 1).astype(bool)
        d = dims[aa]
        v = coords[aa, d]
        c = coords[ii, d]
        bad = np.where(right, c < v, c > v)
        if bad.any():
            k = np.nonzero(bad)[0][0]
            raise AssertionError("node %d violates split of ancestor %d" % (ii[k], aa[k]))
        j += 1
