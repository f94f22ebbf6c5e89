"""cudakdtree_amd -- MI355X (gfx950) k-d tree build and query.

Python binding over the C ABI of ``libcukd_amd.so`` (include/cukd_amd.h).
Device memory and streams come from PyTorch; every call is stream-ordered
on ``torch.cuda.current_stream()`` unless a stream is passed.

The names mirror the reference's header API (cudaKDTree v1.0.1):

  buildTree / buildTree_thrust / buildTree_bitonic / buildTree_inPlace
      -> cukd::buildTree* (builder.h:108-151)
  buildTree_host (host numpy array)    -> builder_thrust.h:114-118
  computeBounds                        -> builder_common.h:59-64
  stackBased.fcp / stackFree.fcp / cct.fcp      -> fcp.h:61-163
  stackBased.knn / stackFree.knn / cct.knn      -> knn.h:83-199
  SpatialKDTree, buildSpatialTree, free         -> spatial-kdtree.h:69-126

Points are a device tensor whose rows are ``layout.stride`` bytes (e.g. a
float32 [n, 3] tensor for float3, or a uint8/int32 view of a payload
struct array).  The library must be built (``python -m
cudakdtree_amd.build``); there is no fallback path.
"""
import ctypes
import math
import os
import types

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libcukd_amd.so")
_lib = None

OK = 0
ERR_INVALID, ERR_HIP, ERR_NO_BOUNDS, ERR_ALLOC, ERR_UNSUPPORTED = -1, -2, -3, -4, -5
SCALAR_F32, SCALAR_I32 = 0, 1
BUILDER_ID_DEFAULT, BUILDER_ID_THRUST, BUILDER_ID_BITONIC, BUILDER_ID_INPLACE = 0, 1, 2, 3
BUILDER_DEFAULT, BUILDER_THRUST = BUILDER_ID_DEFAULT, BUILDER_ID_THRUST
BUILDER_BITONIC, BUILDER_INPLACE = BUILDER_ID_BITONIC, BUILDER_ID_INPLACE
STACK_BASED, STACK_FREE, CCT = 0, 1, 2
LIST_FIXED, LIST_HEAP = 0, 1
QUERY_AUTOTUNE = 2
QUERY_SORT = 1
QUERY_ONE_PER_LANE = 4
QUERY_PRESORTED = 8

# every entry point declared in include/cukd_amd.h
EXPORTED_SYMBOLS = (
    "cukd_build_tree", "cukd_build_tree_host", "cukd_build_tree_temp_bytes",
    "cukd_compute_bounds",
    "cukd_fcp", "cukd_knn", "cukd_knn_noalloc", "cukd_spatial_build", "cukd_spatial_free",
    "cukd_spatial_fcp", "cukd_spatial_knn", "cukd_set_traversal_stats",
    "cukd_knn_accumulate", "cukd_knn_extract", "cukd_query_order",
    "cukd_forest_create", "cukd_forest_workspace_bytes", "cukd_forest_build",
    "cukd_forest_destroy", "cukd_forest_fcp", "cukd_forest_knn", "cukd_forest_reserve_queries",
    "cukd_server_create", "cukd_server_submit", "cukd_server_poll",
    "cukd_server_get_stats", "cukd_server_destroy",
    "cukd_last_error", "cukd_version",
)


class CukdError(RuntimeError):
    """Raised for a non-zero status of a C-ABI call (std::runtime_error in
    the reference, common.h:276-287)."""

    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


class Layout(ctypes.Structure):
    """Runtime data_traits: see cukd_layout in include/cukd_amd.h."""
    _fields_ = [("num_dims", ctypes.c_int32), ("scalar_type", ctypes.c_int32),
                ("stride", ctypes.c_int32), ("point_offset", ctypes.c_int32),
                ("dim_offset", ctypes.c_int32), ("dim_bytes", ctypes.c_int32)]

    def __init__(self, num_dims, scalar_type=SCALAR_F32, stride=None,
                 point_offset=0, dim_offset=-1, dim_bytes=4):
        super().__init__(num_dims, scalar_type,
                         stride if stride is not None else 4 * num_dims,
                         point_offset, dim_offset, dim_bytes)

    @property
    def explicit_dim(self):
        return self.dim_offset >= 0

    @property
    def scalar_dtype(self):
        return np.float32 if self.scalar_type == SCALAR_F32 else np.int32

    def __repr__(self):
        return ("Layout(num_dims=%d, scalar_type=%d, stride=%d, point_offset=%d, "
                "dim_offset=%d, dim_bytes=%d)" % (self.num_dims, self.scalar_type,
                                                  self.stride, self.point_offset,
                                                  self.dim_offset, self.dim_bytes))


def float_points(d):
    """float2/float3/float4 / vec_float<d> (cukd-math.h:43-181)."""
    return Layout(d, SCALAR_F32)


def int_points(d):
    """int2/int3/int4 (cukd-math.h:50-52)."""
    return Layout(d, SCALAR_I32)


def point_plus_payload(d, scalar_type=SCALAR_F32, explicit_dim=False):
    """{vecN pos; int payload; int splitDim} as in
    testing/testBuildersSameResult.cu:188-193 (splitDim used only when
    explicit_dim)."""
    return Layout(d, scalar_type, 4 * d + 8, 0, 4 * d + 4 if explicit_dim else -1, 4)


class MemResource(ctypes.Structure):
    """cukd_mem_resource: user allocation callbacks (GpuMemoryResource)."""
    _fields_ = [("ctx", ctypes.c_void_p),
                ("malloc_fn", ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p,
                                               ctypes.POINTER(ctypes.c_void_p),
                                               ctypes.c_size_t, ctypes.c_void_p)),
                ("free_fn", ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p))]


class SpatialNode(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_float), ("offset", ctypes.c_uint32),
                ("count", ctypes.c_uint16), ("dim", ctypes.c_int16)]


class SpatialTree(ctypes.Structure):
    """cukd_spatial_tree (spatial-kdtree.h:69-98)."""
    _fields_ = [("bounds", ctypes.c_float * 16), ("nodes", ctypes.c_void_p),
                ("prim_ids", ctypes.c_void_p), ("data", ctypes.c_void_p),
                ("num_prims", ctypes.c_int32), ("num_nodes", ctypes.c_int32)]


def lib():
    """Loads libcukd_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise RuntimeError("libcukd_amd.so is not built: run `python -m cudakdtree_amd.build`")
    L = ctypes.CDLL(_LIB_PATH)
    P, i32, f32 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_float
    LP = ctypes.POINTER(Layout)
    L.cukd_build_tree.argtypes = [P, i32, LP, ctypes.c_int, P, P, P]
    L.cukd_build_tree_host.argtypes = [P, i32, LP, P]
    L.cukd_build_tree_temp_bytes.argtypes = [i32, LP, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_size_t)]
    L.cukd_compute_bounds.argtypes = [P, P, i32, LP, P]
    L.cukd_fcp.argtypes = [ctypes.c_int, P, i32, LP, P, P, i32, f32, f32, P, P,
                           ctypes.c_int, P, P]
    L.cukd_knn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, i32, LP, P, P,
                           i32, f32, P, P, P, ctypes.c_int, P, P]
    L.cukd_knn_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, i32, LP, P, P,
                                   i32, f32, P, P, P, ctypes.c_int, P]
    SP = ctypes.POINTER(SpatialTree)
    L.cukd_spatial_build.argtypes = [SP, P, i32, LP, ctypes.c_int, P, P]
    L.cukd_spatial_free.argtypes = [SP, P, P]
    L.cukd_spatial_fcp.argtypes = [ctypes.c_int, SP, LP, P, i32, f32, P, P, P]
    L.cukd_spatial_knn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, SP, LP, P,
                                   i32, f32, P, P, P, P]
    L.cukd_knn_accumulate.argtypes = [ctypes.c_int, ctypes.c_int, P, i32, LP, P, P, i32, P,
                                       ctypes.c_int, f32, i32, ctypes.c_int, P, P]
    L.cukd_knn_extract.argtypes = [P, i32, ctypes.c_int, P, P, P, P]
    L.cukd_query_order.argtypes = [P, i32, i32, P, P, P]
    L.cukd_forest_create.argtypes = [ctypes.POINTER(P), P, i32, LP, P, P]
    L.cukd_forest_workspace_bytes.argtypes = [P, ctypes.POINTER(ctypes.c_size_t)]
    L.cukd_forest_build.argtypes = [P, P, P, P]
    L.cukd_forest_destroy.argtypes = [P, P]
    L.cukd_forest_fcp.argtypes = [ctypes.c_int, P, P, P, P, i32, i32, P, P, f32, f32, P, P,
                                  ctypes.c_int, P]
    L.cukd_forest_knn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, P, i32, i32,
                                  P, P, f32, P, P, P, ctypes.c_int, P]
    L.cukd_forest_reserve_queries.argtypes = [P, i32, P]
    L.cukd_server_create.argtypes = [ctypes.POINTER(P), P, P, P, ctypes.c_int, ctypes.c_size_t,
                                     i32, P]
    L.cukd_server_submit.argtypes = [P, i32, P, i32, ctypes.c_int, f32, P, P,
                                     ctypes.POINTER(ctypes.c_uint64)]
    L.cukd_server_poll.argtypes = [P, ctypes.POINTER(Completion), i32,
                                   ctypes.POINTER(ctypes.c_int32), i32]
    L.cukd_server_get_stats.argtypes = [P, ctypes.POINTER(ServerStats)]
    L.cukd_server_destroy.argtypes = [P]
    L.cukd_set_traversal_stats.argtypes = [P]
    L.cukd_last_error.restype = ctypes.c_char_p
    L.cukd_last_error.argtypes = []
    L.cukd_version.restype = ctypes.c_int
    _lib = L
    return L


def _check(rc):
    if rc != OK:
        raise CukdError(rc, lib().cukd_last_error().decode())


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _rows(points, layout):
    """Number of elements in a device tensor of layout.stride-byte rows."""
    nbytes = points.numel() * points.element_size()
    if nbytes % layout.stride:
        raise ValueError("tensor size is not a multiple of the element stride")
    if not points.is_contiguous():
        raise ValueError("points must be contiguous")
    return nbytes // layout.stride


# ----------------------------------------------------------------------------
# builders (builder.h:108-151)
# ----------------------------------------------------------------------------

def build_tree(points, layout, builder=BUILDER_DEFAULT, world_bounds=None,
               stream=None, n=None):
    """Builds the left-balanced tree in place over ``points``."""
    n = _rows(points, layout) if n is None else n
    _check(lib().cukd_build_tree(_ptr(points), n, ctypes.byref(layout), builder,
                                 _ptr(world_bounds), _stream(stream), None))
    return points


def buildTree(points, layout, world_bounds=None, stream=None):
    return build_tree(points, layout, BUILDER_DEFAULT, world_bounds, stream)


def buildTree_thrust(points, layout, world_bounds=None, stream=None):
    return build_tree(points, layout, BUILDER_THRUST, world_bounds, stream)


def buildTree_bitonic(points, layout, world_bounds=None, stream=None):
    return build_tree(points, layout, BUILDER_BITONIC, world_bounds, stream)


def buildTree_inPlace(points, layout, world_bounds=None, stream=None):
    return build_tree(points, layout, BUILDER_INPLACE, world_bounds, stream)


buildTree_inplace = buildTree_inPlace


def buildTree_host(points, layout, want_bounds=False):
    """buildTree_host (builder_thrust.h:114-118): builds the tree in place
    over a host numpy array of ``layout.stride``-byte rows (C-contiguous,
    writeable).  Returns the host bounds box (lower[D] then upper[D]) when
    ``want_bounds``, else None."""
    if not (isinstance(points, np.ndarray) and points.flags.c_contiguous
            and points.flags.writeable):
        raise ValueError("points must be a writeable C-contiguous numpy array")
    if points.nbytes % layout.stride:
        raise ValueError("array size is not a multiple of the element stride")
    box = np.zeros(2 * layout.num_dims, dtype=layout.scalar_dtype) if want_bounds else None
    _check(lib().cukd_build_tree_host(points.ctypes.data_as(ctypes.c_void_p),
                                      points.nbytes // layout.stride, ctypes.byref(layout),
                                      None if box is None else box.ctypes.data_as(
                                          ctypes.c_void_p)))
    return box


def build_temp_bytes(n, layout, builder=BUILDER_DEFAULT):
    out = ctypes.c_size_t(0)
    _check(lib().cukd_build_tree_temp_bytes(n, ctypes.byref(layout), builder,
                                            ctypes.byref(out)))
    return out.value


def bounds_tensor(layout, device="cuda"):
    """A device buffer for box_t<point_t>: lower[D] then upper[D]."""
    import torch
    dt = torch.float32 if layout.scalar_type == SCALAR_F32 else torch.int32
    return torch.empty(2 * layout.num_dims, dtype=dt, device=device)


def computeBounds(bounds, points, layout, stream=None):
    n = _rows(points, layout)
    _check(lib().cukd_compute_bounds(_ptr(bounds), _ptr(points), n, ctypes.byref(layout),
                                     _stream(stream)))
    return bounds


# ----------------------------------------------------------------------------
# queries (fcp.h, knn.h)
# ----------------------------------------------------------------------------

def fcp(tree, layout, queries, traversal=STACK_BASED, cutoff=math.inf, eps=0.0,
        world_bounds=None, flags=0, stream=None, out=None):
    """Batched FCP; returns (ids int32[nq], dist2 float32[nq])."""
    import torch
    n = _rows(tree, layout)
    nq = queries.shape[0]
    if out is None:
        ids = torch.empty(nq, dtype=torch.int32, device=queries.device)
        d2 = torch.empty(nq, dtype=torch.float32, device=queries.device)
    else:
        ids, d2 = out
    _check(lib().cukd_fcp(traversal, _ptr(tree), n, ctypes.byref(layout), _ptr(world_bounds),
                          _ptr(queries), nq, cutoff, eps, _ptr(ids), _ptr(d2), flags,
                          _stream(stream), None))
    return ids, d2


def knn(tree, layout, queries, k, traversal=STACK_BASED, list_kind=LIST_FIXED,
        cutoff=math.inf, world_bounds=None, flags=0, stream=None, out=None):
    """Batched kNN; returns (ids int32[nq,k], dist2 float32[nq,k],
    max_radius2 float32[nq])."""
    import torch
    n = _rows(tree, layout)
    nq = queries.shape[0]
    if out is None:
        ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
        d2 = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
        ret = torch.empty(nq, dtype=torch.float32, device=queries.device)
    else:
        ids, d2, ret = out
    _check(lib().cukd_knn(traversal, list_kind, k, _ptr(tree), n, ctypes.byref(layout),
                          _ptr(world_bounds), _ptr(queries), nq, cutoff, _ptr(ids),
                          _ptr(d2), _ptr(ret), flags, _stream(stream), None))
    return ids, d2, ret


def query_order(queries, stream=None):
    """The Morton launch order CUKD_QUERY_SORT uses (int32 permutation)."""
    import torch
    nq, D = queries.shape
    order = torch.empty(nq, dtype=torch.int32, device=queries.device)
    _check(lib().cukd_query_order(_ptr(queries), nq, D, _ptr(order), _stream(stream), None))
    return order


def knn_lists(nq, k, device="cuda"):
    """Persistent per-query candidate lists for knn_accumulate (uint64 as
    int64 storage, nq x k)."""
    import torch
    return torch.empty((nq, k), dtype=torch.int64, device=device)


def knn_accumulate(tree, layout, queries, lists, init, traversal=STACK_FREE,
                   cutoff=math.inf, id_offset=0, world_bounds=None, flags=0, stream=None):
    """One round of a multi-round kNN (samples/mpiHugeQuery.cu:75-88): each
    query continues its FlexHeapCandidateList in ``lists`` (nq x k)."""
    n = _rows(tree, layout)
    nq, k = lists.shape
    assert queries.shape[0] == nq
    _check(lib().cukd_knn_accumulate(traversal, k, _ptr(tree), n, ctypes.byref(layout),
                                     _ptr(world_bounds), _ptr(queries), nq, _ptr(lists),
                                     1 if init else 0, cutoff, id_offset, flags,
                                     _stream(stream), None))
    return lists


def knn_extract(lists, stream=None):
    """extractFinalResult (samples/mpiHugeQuery.cu:90-103): returns
    (ids int32[nq,k] and dist2 float32[nq,k] ascending, radius float32[nq])."""
    import torch
    nq, k = lists.shape
    ids = torch.empty((nq, k), dtype=torch.int32, device=lists.device)
    d2 = torch.empty((nq, k), dtype=torch.float32, device=lists.device)
    rad = torch.empty(nq, dtype=torch.float32, device=lists.device)
    _check(lib().cukd_knn_extract(_ptr(lists), nq, k, _ptr(ids), _ptr(d2), _ptr(rad),
                                  _stream(stream)))
    return ids, d2, rad


class Completion(ctypes.Structure):
    """cukd_completion"""
    _fields_ = [("ticket", ctypes.c_uint64), ("status", ctypes.c_int32),
                ("num_queries", ctypes.c_int32), ("queue_ms", ctypes.c_float),
                ("total_ms", ctypes.c_float)]


class ServerStats(ctypes.Structure):
    """cukd_server_stats"""
    _fields_ = [("requests_submitted", ctypes.c_uint64), ("requests_done", ctypes.c_uint64),
                ("queries_done", ctypes.c_uint64), ("batches", ctypes.c_uint64)]


class Server:
    """Serving loop over a built forest (cukd_server_*).  Requests carry host
    (numpy) queries; results land in numpy arrays owned by the request."""

    def __init__(self, forest, points, world_bounds=None, traversal=STACK_BASED,
                 budget_bytes=256 << 20, max_requests_per_batch=0, stream=None):
        self.forest = forest
        self._keep = {}
        self._h = ctypes.c_void_p()
        self._refs = (points, world_bounds)
        _check(lib().cukd_server_create(ctypes.byref(self._h), forest._h, _ptr(points),
                                        _ptr(world_bounds), traversal, budget_bytes,
                                        max_requests_per_batch, _stream(stream)))

    def submit(self, tree, queries, k=0, cutoff=math.inf, out=None):
        """queries: float32 numpy (nq, D).  Returns the ticket; results are
        in result(ticket) once polled complete.  out: optional (ids, dist2)
        host arrays of nq*max(1, k) int32 / float32 elements to receive the
        results (reused buffers spare the page faults of fresh ones)."""
        import numpy as np
        q = np.ascontiguousarray(queries, dtype=np.float32)
        nq = len(q)
        w = max(1, k)
        if out is None:
            ids = np.empty((nq, w) if k else nq, np.int32)
            d2 = np.empty((nq, w) if k else nq, np.float32)
        else:
            ids, d2 = out
            if (ids.dtype != np.int32 or d2.dtype != np.float32 or ids.size != nq * w
                    or d2.size != nq * w or not ids.flags.c_contiguous
                    or not d2.flags.c_contiguous):
                raise ValueError("out must be C-contiguous int32/float32 arrays of "
                                 "nq*max(1,k) elements")
        t = ctypes.c_uint64(0)
        _check(lib().cukd_server_submit(self._h, tree, q.ctypes.data, nq, k, cutoff,
                                        ids.ctypes.data, d2.ctypes.data, ctypes.byref(t)))
        self._keep[t.value] = (q, ids, d2)
        return t.value

    def poll(self, wait_for=0, max_items=4096):
        """List of completed Completion records (waits for wait_for)."""
        buf = (Completion * max_items)()
        cnt = ctypes.c_int32(0)
        _check(lib().cukd_server_poll(self._h, buf, max_items, ctypes.byref(cnt), wait_for))
        return [buf[i] for i in range(cnt.value)]

    def result(self, ticket):
        """(ids, dist2) of a completed request; releases its buffers."""
        q, ids, d2 = self._keep.pop(ticket)
        return ids, d2

    def stats(self):
        s = ServerStats()
        _check(lib().cukd_server_get_stats(self._h, ctypes.byref(s)))
        return s

    def close(self):
        if self._h:
            rc = lib().cukd_server_destroy(self._h)
            self._h = ctypes.c_void_p()
            _check(rc)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Forest:
    """Many independent trees in one array, built and queried together
    (cukd_forest_*).  ``offsets``: num_trees+1 element offsets (host)."""

    def __init__(self, offsets, layout, stream=None):
        import numpy as np
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        self.offsets = off
        self.num_trees = len(off) - 1
        self.layout = layout
        self._h = ctypes.c_void_p()
        _check(lib().cukd_forest_create(ctypes.byref(self._h), off.ctypes.data, self.num_trees,
                                        ctypes.byref(layout), _stream(stream), None))

    @property
    def workspace_bytes(self):
        out = ctypes.c_size_t(0)
        _check(lib().cukd_forest_workspace_bytes(self._h, ctypes.byref(out)))
        return out.value

    def bounds_tensor(self, device="cuda"):
        import torch
        dt = torch.float32 if self.layout.scalar_type == SCALAR_F32 else torch.int32
        return torch.empty((self.num_trees, 2 * self.layout.num_dims), dtype=dt, device=device)

    def build(self, points, world_bounds=None, stream=None):
        _check(lib().cukd_forest_build(self._h, _ptr(points), _ptr(world_bounds),
                                       _stream(stream)))
        return points

    def reserve_queries(self, max_queries, stream=None):
        _check(lib().cukd_forest_reserve_queries(self._h, max_queries, _stream(stream)))

    def fcp(self, points, queries, seg_offsets, seg_tree, traversal=STACK_BASED,
            cutoff=math.inf, eps=0.0, world_bounds=None, flags=0, stream=None, out=None):
        import torch
        nq = queries.shape[0]
        if out is None:
            ids = torch.empty(nq, dtype=torch.int32, device=queries.device)
            d2 = torch.empty(nq, dtype=torch.float32, device=queries.device)
        else:
            ids, d2 = out
        _check(lib().cukd_forest_fcp(traversal, self._h, _ptr(points), _ptr(world_bounds),
                                     _ptr(queries), nq, len(seg_tree), _ptr(seg_offsets),
                                     _ptr(seg_tree), cutoff, eps, _ptr(ids), _ptr(d2), flags,
                                     _stream(stream)))
        return ids, d2

    def knn(self, points, queries, k, seg_offsets, seg_tree, traversal=STACK_BASED,
            list_kind=LIST_FIXED, cutoff=math.inf, world_bounds=None, flags=0, stream=None,
            out=None):
        import torch
        nq = queries.shape[0]
        if out is None:
            ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
            d2 = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
            ret = torch.empty(nq, dtype=torch.float32, device=queries.device)
        else:
            ids, d2, ret = out
        _check(lib().cukd_forest_knn(traversal, list_kind, k, self._h, _ptr(points),
                                     _ptr(world_bounds), _ptr(queries), nq, len(seg_tree),
                                     _ptr(seg_offsets), _ptr(seg_tree), cutoff, _ptr(ids),
                                     _ptr(d2), _ptr(ret), flags, _stream(stream)))
        return ids, d2, ret

    def close(self, stream=None):
        if self._h:
            _check(lib().cukd_forest_destroy(self._h, _stream(stream)))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def set_traversal_stats(counter):
    """counter: a device int64 tensor of one element, or None."""
    _check(lib().cukd_set_traversal_stats(_ptr(counter)))


def _ns(traversal):
    def _fcp(tree, layout, queries, cutoff=math.inf, eps=0.0, world_bounds=None, **kw):
        return fcp(tree, layout, queries, traversal, cutoff, eps, world_bounds, **kw)

    def _knn(tree, layout, queries, k, list_kind=LIST_FIXED, cutoff=math.inf,
             world_bounds=None, **kw):
        return knn(tree, layout, queries, k, traversal, list_kind, cutoff, world_bounds, **kw)
    return types.SimpleNamespace(fcp=_fcp, knn=_knn)


stackBased = _ns(STACK_BASED)
stackFree = _ns(STACK_FREE)
cct = _ns(CCT)


# ----------------------------------------------------------------------------
# spatial k-d tree (spatial-kdtree.h:69-126)
# ----------------------------------------------------------------------------

class SpatialKDTree:
    """Device-resident spatial k-d tree; ``bounds`` is a host list of
    2*num_dims floats.  Keeps the point tensor alive (the tree references
    it, the reference's SpatialKDTree::data)."""

    def __init__(self, raw, points, layout):
        self.raw = raw
        self.points = points
        self.layout = layout

    @property
    def num_nodes(self):
        return self.raw.num_nodes

    @property
    def num_prims(self):
        return self.raw.num_prims

    @property
    def bounds(self):
        """float bounds, or int32 ones for int points"""
        b = np.ctypeslib.as_array(self.raw.bounds)[:2 * self.layout.num_dims].copy()
        return list(b.view(self.layout.scalar_dtype))

    def nodes_numpy(self):
        """Copies the node array to the host as a structured array."""
        import torch
        n = self.raw.num_nodes
        buf = torch.empty(n * 12, dtype=torch.uint8, device=self.points.device)
        _copy_dev(buf, self.raw.nodes, n * 12)
        pos = "<f4" if self.layout.scalar_type == SCALAR_F32 else "<i4"
        dt = np.dtype([("pos", pos), ("offset", "<u4"), ("count", "<u2"), ("dim", "<i2")])
        return buf.cpu().numpy().view(dt)

    def prims_numpy(self):
        import torch
        n = self.raw.num_prims
        buf = torch.empty(max(n, 1), dtype=torch.int32, device=self.points.device)
        if n:
            _copy_dev(buf, self.raw.prim_ids, n * 4)
        return buf.cpu().numpy().view(np.uint32)[:n]


def _copy_dev(dst_tensor, src_ptr, nbytes):
    import torch
    torch.cuda.synchronize()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rc = hip.hipMemcpy(ctypes.c_void_p(dst_tensor.data_ptr()), ctypes.c_void_p(src_ptr), nbytes, 3)
    if rc != 0:
        raise RuntimeError("hipMemcpy failed (%d)" % rc)


def buildSpatialTree(points, layout, leaf_threshold=0, stream=None):
    """cukd::buildTree(SpatialKDTree&, ...): synchronises the stream."""
    n = _rows(points, layout)
    raw = SpatialTree()
    _check(lib().cukd_spatial_build(ctypes.byref(raw), _ptr(points), n, ctypes.byref(layout),
                                    leaf_threshold, _stream(stream), None))
    return SpatialKDTree(raw, points, layout)


def free(tree, stream=None):
    """cukd::free(SpatialKDTree&)."""
    _check(lib().cukd_spatial_free(ctypes.byref(tree.raw), _stream(stream), None))


def spatial_fcp(tree, queries, traversal=STACK_BASED, cutoff=math.inf, stream=None):
    import torch
    nq = queries.shape[0]
    ids = torch.empty(nq, dtype=torch.int32, device=queries.device)
    d2 = torch.empty(nq, dtype=torch.float32, device=queries.device)
    _check(lib().cukd_spatial_fcp(traversal, ctypes.byref(tree.raw), ctypes.byref(tree.layout),
                                  _ptr(queries), nq, cutoff, _ptr(ids), _ptr(d2), _stream(stream)))
    return ids, d2


def spatial_knn(tree, queries, k, traversal=STACK_BASED, list_kind=LIST_FIXED,
                cutoff=math.inf, stream=None):
    import torch
    nq = queries.shape[0]
    ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
    d2 = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
    ret = torch.empty(nq, dtype=torch.float32, device=queries.device)
    _check(lib().cukd_spatial_knn(traversal, list_kind, k, ctypes.byref(tree.raw),
                                  ctypes.byref(tree.layout), _ptr(queries), nq, cutoff, _ptr(ids),
                                  _ptr(d2), _ptr(ret), _stream(stream)))
    return ids, d2, ret
