"""ctypes wrapper over the CPU parity oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY.  Only the test suite, the smoke check and the
``cpu_baseline`` leg of ``bench.py`` import this module; the product
package never does.  ``make -C oracle`` builds the library.

Arrays are numpy arrays; a point set is any C-contiguous array whose rows
are ``layout.stride`` bytes (e.g. ``float32[n, 3]`` for float3, or a
structured dtype for payload types).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

STACK_BASED, STACK_FREE, CCT = 0, 1, 2
LIST_FIXED, LIST_HEAP = 0, 1
F32, I32 = 0, 1


class OcLayout(ctypes.Structure):
    _fields_ = [("num_dims", ctypes.c_int32), ("scalar_type", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("point_offset", ctypes.c_int32),
                ("dim_offset", ctypes.c_int32), ("dim_bytes", ctypes.c_int32)]


class OcSpatialNode(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float), ("offset", ctypes.c_uint32),
                ("count", ctypes.c_uint16), ("dim", ctypes.c_int16)]


SPATIAL_NODE_DTYPE = np.dtype([("pos", "<f4"), ("offset", "<u4"),
                               ("count", "<u2"), ("dim", "<i2")])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64, i32, f32 = ctypes.c_int64, ctypes.c_int, ctypes.c_float
        L.oc_level_of.argtypes = [i64]; L.oc_level_of.restype = i32
        L.oc_num_levels_for.argtypes = [i64]; L.oc_num_levels_for.restype = i32
        L.oc_subtree_size.argtypes = [i64, i64]; L.oc_subtree_size.restype = i64
        L.oc_segment_begin.argtypes = [i32, i64, i64]; L.oc_segment_begin.restype = i64
        L.oc_pivot_pos.argtypes = [i32, i64, i64]; L.oc_pivot_pos.restype = i64
        L.oc_compute_bounds.argtypes = [P, i64, P, P]; L.oc_compute_bounds.restype = None
        L.oc_widest_dim.argtypes = [P, P]; L.oc_widest_dim.restype = i32
        L.oc_build_host.argtypes = [P, i64, P, P]; L.oc_build_host.restype = i32
        L.oc_fcp.argtypes = [i32, P, i64, P, P, P, i64, f32, f32, P, P, P, i32]
        L.oc_fcp.restype = None
        L.oc_knn.argtypes = [i32, i32, i32, P, i64, P, P, P, i64, f32, P, P, P, P, i32]
        L.oc_knn.restype = None
        L.oc_knn_accumulate.argtypes = [i32, i32, P, i64, P, P, P, i64, P, i32, f32, i32, P,
                                        i32]
        L.oc_knn_accumulate.restype = None
        L.oc_knn_extract.argtypes = [P, i64, i32, P, P, P]
        L.oc_knn_extract.restype = None
        L.oc_spatial_build.argtypes = [P, i64, P, i32, P, P, P]
        L.oc_spatial_build.restype = i64
        L.oc_spatial_fcp.argtypes = [i32, P, P, P, P, P, P, i64, f32, P, P, i32]
        L.oc_spatial_fcp.restype = None
        L.oc_spatial_knn.argtypes = [i32, i32, i32, P, P, P, P, P, P, i64, f32, P, P, P, i32]
        L.oc_spatial_knn.restype = None
        L.oc_hash_words.argtypes = [P, i64]; L.oc_hash_words.restype = ctypes.c_uint64
        _LIB = L
    return _LIB


def _layout(layout):
    return OcLayout(layout.num_dims, layout.scalar_type, layout.stride,
                    layout.point_offset, layout.dim_offset, layout.dim_bytes)


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _scalar_dtype(layout):
    return np.float32 if layout.scalar_type == F32 else np.int32


def _nthreads(n):
    return n if n else (os.cpu_count() or 1)


def level_of(node):
    return lib().oc_level_of(node)


def num_levels_for(n):
    return lib().oc_num_levels_for(n)


def subtree_size(node, n):
    return lib().oc_subtree_size(node, n)


def segment_begin(step, n, subtree):
    return lib().oc_segment_begin(step, n, subtree)


def pivot_pos(step, n, subtree):
    return lib().oc_pivot_pos(step, n, subtree)


def compute_bounds(points, layout):
    pts = np.ascontiguousarray(points)
    box = np.zeros(2 * layout.num_dims, dtype=_scalar_dtype(layout))
    lay = _layout(layout)
    lib().oc_compute_bounds(_ptr(pts), len(pts), ctypes.byref(lay), _ptr(box))
    return box


def widest_dim(box, layout):
    box = np.ascontiguousarray(box, dtype=_scalar_dtype(layout))
    lay = _layout(layout)
    return lib().oc_widest_dim(_ptr(box), ctypes.byref(lay))


def build_host(points, layout, want_bounds=False):
    """buildTree_host restatement; returns (tree, bounds or None)."""
    tree = np.array(points, copy=True, order="C")
    box = np.zeros(2 * layout.num_dims, dtype=_scalar_dtype(layout)) \
        if (want_bounds or layout.dim_offset >= 0) else None
    lay = _layout(layout)
    rc = lib().oc_build_host(_ptr(tree), len(tree), ctypes.byref(lay), _ptr(box))
    if rc != 0:
        raise RuntimeError("oc_build_host failed (%d)" % rc)
    return tree, box


def fcp(method, tree, layout, queries, cutoff=float("inf"), eps=0.0,
        world_bounds=None, threads=0, want_visits=False):
    tree = np.ascontiguousarray(tree)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = len(q)
    ids = np.empty(nq, np.int32)
    d2 = np.empty(nq, np.float32)
    vis = np.empty(nq, np.int64) if want_visits else None
    wb = None if world_bounds is None else np.ascontiguousarray(world_bounds, np.float32)
    lay = _layout(layout)
    lib().oc_fcp(method, _ptr(tree), len(tree), ctypes.byref(lay), _ptr(wb), _ptr(q),
                 nq, cutoff, eps, _ptr(ids), _ptr(d2), _ptr(vis), _nthreads(threads))
    return (ids, d2, vis) if want_visits else (ids, d2)


def knn(method, list_kind, k, tree, layout, queries, cutoff=float("inf"),
        world_bounds=None, threads=0, want_visits=False):
    tree = np.ascontiguousarray(tree)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = len(q)
    ids = np.empty((nq, k), np.int32)
    d2 = np.empty((nq, k), np.float32)
    ret = np.empty(nq, np.float32)
    vis = np.empty(nq, np.int64) if want_visits else None
    wb = None if world_bounds is None else np.ascontiguousarray(world_bounds, np.float32)
    lay = _layout(layout)
    lib().oc_knn(method, list_kind, k, _ptr(tree), len(tree), ctypes.byref(lay),
                 _ptr(wb), _ptr(q), nq, cutoff, _ptr(ids), _ptr(d2), _ptr(ret),
                 _ptr(vis), _nthreads(threads))
    return (ids, d2, ret, vis) if want_visits else (ids, d2, ret)


def knn_accumulate(method, k, tree, layout, queries, lists, init, cutoff=float("inf"),
                   id_offset=0, world_bounds=None, threads=0):
    """One round of kNN into persistent FlexHeap lists (uint64 (nq, k),
    updated in place); see oc_knn_accumulate."""
    tree = np.ascontiguousarray(tree)
    q = np.ascontiguousarray(queries, dtype=np.float32)
    assert lists.dtype == np.uint64 and lists.flags.c_contiguous and lists.shape == (len(q), k)
    wb = None if world_bounds is None else np.ascontiguousarray(world_bounds, np.float32)
    lay = _layout(layout)
    lib().oc_knn_accumulate(method, k, _ptr(tree), len(tree), ctypes.byref(lay), _ptr(wb),
                            _ptr(q), len(q), _ptr(lists), 1 if init else 0, cutoff, id_offset,
                            None, _nthreads(threads))
    return lists


def knn_extract(lists):
    """(ids, dist2) ascending per query and the extractFinalResult radius."""
    lists = np.ascontiguousarray(lists, dtype=np.uint64)
    nq, k = lists.shape
    ids = np.empty((nq, k), np.int32)
    d2 = np.empty((nq, k), np.float32)
    rad = np.empty(nq, np.float32)
    lib().oc_knn_extract(_ptr(lists), nq, k, _ptr(ids), _ptr(d2), _ptr(rad))
    return ids, d2, rad


def spatial_build(points, layout, leaf_threshold=0):
    """Returns (nodes[structured], prim_ids uint32, bounds float32[2D])."""
    pts = np.ascontiguousarray(points)
    lay = _layout(layout)
    n = len(pts)
    num = lib().oc_spatial_build(_ptr(pts), n, ctypes.byref(lay), leaf_threshold,
                                 None, None, None)
    nodes = np.zeros(num, SPATIAL_NODE_DTYPE)
    prims = np.zeros(max(n, 1), np.uint32)
    # box_t<point_t>: float, or int32 for int points (pos likewise holds
    # int32 bits then)
    bounds = np.zeros(2 * layout.num_dims, _scalar_dtype(layout))
    lib().oc_spatial_build(_ptr(pts), n, ctypes.byref(lay), leaf_threshold,
                           _ptr(nodes), _ptr(prims), _ptr(bounds))
    return nodes, prims[:n], bounds


def spatial_fcp(method, nodes, prims, bounds, points, layout, queries,
                cutoff=float("inf"), threads=0):
    q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = len(q)
    ids = np.empty(nq, np.int32)
    d2 = np.empty(nq, np.float32)
    lay = _layout(layout)
    pr = np.ascontiguousarray(prims) if len(prims) else np.zeros(1, np.uint32)
    lib().oc_spatial_fcp(method, _ptr(nodes), _ptr(pr), _ptr(bounds),
                         _ptr(np.ascontiguousarray(points)), ctypes.byref(lay),
                         _ptr(q), nq, cutoff, _ptr(ids), _ptr(d2), _nthreads(threads))
    return ids, d2


def spatial_knn(method, list_kind, k, nodes, prims, bounds, points, layout,
                queries, cutoff=float("inf"), threads=0):
    q = np.ascontiguousarray(queries, dtype=np.float32)
    nq = len(q)
    ids = np.empty((nq, k), np.int32)
    d2 = np.empty((nq, k), np.float32)
    ret = np.empty(nq, np.float32)
    lay = _layout(layout)
    pr = np.ascontiguousarray(prims) if len(prims) else np.zeros(1, np.uint32)
    lib().oc_spatial_knn(method, list_kind, k, _ptr(nodes), _ptr(pr), _ptr(bounds),
                         _ptr(np.ascontiguousarray(points)), ctypes.byref(lay),
                         _ptr(q), nq, cutoff, _ptr(ids), _ptr(d2), _ptr(ret),
                         _nthreads(threads))
    return ids, d2, ret


def hash_words(arr):
    a = np.ascontiguousarray(arr)
    return int(lib().oc_hash_words(_ptr(a), a.nbytes))
