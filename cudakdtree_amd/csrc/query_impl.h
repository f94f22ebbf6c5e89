// query_impl.h -- batched FCP / kNN kernels over a left-balanced k-d tree.
//
// One query per lane.  The traversals and result types are the public
// device-inline ones of include/cukd/ (traverse.h, fcp.h, knn.h), run
// through an accessor over the run-time element layout, so the batched
// path and user kernels share one implementation; ids and dist2 bits
// equal the CPU oracle's, ties included.
#pragma once
#pragma clang fp contract(off)

#include <new>

#include "cukd/knn.h"
#include "cukd_common.h"

namespace cukd_amd {

// node accessor over a cukd_layout-described array
struct TreeView {
  const uint8_t *base;
  int n;
  int stride;
  int point_offset;
  int dim_offset;
  int dim_bytes;
  int num_dims;

  __device__ inline int size() const { return n; }
  template <int D>
  __device__ inline void point(int i, float (&p)[D]) const
  {
    const float *src = (const float *)(base + (size_t)i * stride + point_offset);
#pragma unroll
    for (int d = 0; d < D; d++) p[d] = src[d];
  }
  __device__ inline int dim(int i) const
  {
    if (dim_offset < 0) {
      const int l = 31 - __clz(i + 1);
      return (num_dims & (num_dims - 1)) == 0 ? (l & (num_dims - 1)) : l % num_dims;
    }
    const uint8_t *p = base + (size_t)i * stride + dim_offset;
    if (dim_bytes == 1) return *p;
    if (dim_bytes == 2) return *(const uint16_t *)p;
    return *(const int32_t *)p;
  }
};

// TreeView with the query's dimension count known at compile time: the
// round-robin split dim of a node is then a constant-divisor modulo
template <int D>
struct TreeViewD : TreeView {
  __device__ inline int dim(int i) const
  {
    if (dim_offset >= 0) return TreeView::dim(i);
    const int l = 31 - __clz(i + 1);
    return (D & (D - 1)) == 0 ? (l & (D - 1)) : l % D;
  }
};

// FixedCandidateList with a run-time k (<= 64); entries live in scratch
struct DynFixedList {
  uint64_t entry[64];
  int k;
  __device__ inline float maxRadius2() const { return __uint_as_float((uint32_t)(entry[k - 1] >> 32)); }
  __device__ inline float initialCullDist2() const { return maxRadius2(); }
  __device__ inline float processCandidate(int id, float d2)
  {
    uint64_t v = ((uint64_t)__float_as_uint(d2) << 32) | (uint32_t)id;
    for (int i = 0; i < k; i++) {
      const uint64_t lo = entry[i] < v ? entry[i] : v;
      const uint64_t hi = entry[i] < v ? v : entry[i];
      entry[i] = lo;
      v = hi;
    }
    return maxRadius2();
  }
};

// persistent per-query candidate lists in HBM (multi-round kNN)
struct ListArgs {
  uint64_t *lists;  // nq * k entries, query q at lists[q*k]
  int init;         // fill with (cutoff^2, -1) before the traversal
  int id_offset;    // added to every candidate id
};

// The reference's global-memory FlexHeapCandidateList of a multi-round
// query (samples/mpiHugeQuery.cu:75-88, knn.h:360-421): the k entries are
// loaded from HBM into per-lane storage, updated with the same heap push,
// and stored back, so they end as the reference's in-place list would.
struct AccumList {
  uint64_t mem[64];
  int k;
  int id_offset;
  __device__ inline float maxRadius2() const { return __uint_as_float((uint32_t)(mem[0] >> 32)); }
  __device__ inline float initialCullDist2() const { return maxRadius2(); }
  __device__ inline float processCandidate(int id, float d2)
  {
    cukd::detail::heap_push(mem, k, ((uint64_t)__float_as_uint(d2) << 32) | (uint32_t)(id + id_offset));
    return maxRadius2();
  }
};

// ---------------------------------------------------------------------------
// per-result-type set-up and output
// ---------------------------------------------------------------------------
template <class R> struct ResultOps;

template <> struct ResultOps<AccumList> {
  struct State { AccumList r; };
  __device__ static inline AccumList &init(State &s, float cutoff, int k, int q, const ListArgs &la)
  {
    s.r.k = k;
    s.r.id_offset = la.id_offset;
    const uint64_t *src = la.lists + (size_t)q * k;
    const uint64_t v = ((uint64_t)__float_as_uint(cutoff * cutoff) << 32) | 0xffffffffu;
    for (int i = 0; i < k; i++) s.r.mem[i] = la.init ? v : src[i];
    return s.r;
  }
  __device__ static inline void write(const State &s, int q, int k, int32_t *, float *, float *,
                                      const ListArgs &la)
  {
    uint64_t *dst = la.lists + (size_t)q * k;
    for (int i = 0; i < k; i++) dst[i] = s.r.mem[i];
  }
};

template <> struct ResultOps<cukd::FCPResult> {
  struct State { cukd::FCPResult r; };
  __device__ static inline cukd::FCPResult &init(State &s, float cutoff, int, int, const ListArgs &)
  {
    s.r.clear(cutoff * cutoff);
    return s.r;
  }
  __device__ static inline void write(const State &s, int q, int, int32_t *ids, float *dist2, float *,
                                      const ListArgs &)
  {
    ids[q] = s.r.closestPrimID;
    if (dist2) dist2[q] = s.r.closestDist2;
  }
};

template <int K> struct ResultOps<cukd::FixedCandidateList<K>> {
  struct State { cukd::FixedCandidateList<K> r; __device__ State() : r(0.f) {} };
  __device__ static inline cukd::FixedCandidateList<K> &init(State &s, float cutoff, int, int,
                                                             const ListArgs &)
  {
    s.r = cukd::FixedCandidateList<K>(cutoff);
    return s.r;
  }
  __device__ static inline void write(const State &s, int q, int, int32_t *ids, float *dist2, float *ret,
                                      const ListArgs &)
  {
#pragma unroll
    for (int i = 0; i < K; i++) {
      ids[(size_t)q * K + i] = (int32_t)(uint32_t)s.r.entry[i];
      dist2[(size_t)q * K + i] = __uint_as_float((uint32_t)(s.r.entry[i] >> 32));
    }
    if (ret) ret[q] = s.r.maxRadius2();
  }
};

template <> struct ResultOps<DynFixedList> {
  struct State { DynFixedList r; };
  __device__ static inline DynFixedList &init(State &s, float cutoff, int k, int, const ListArgs &)
  {
    s.r.k = k;
    const uint64_t v = ((uint64_t)__float_as_uint(cutoff * cutoff) << 32) | 0xffffffffu;
    for (int i = 0; i < k; i++) s.r.entry[i] = v;
    return s.r;
  }
  __device__ static inline void write(const State &s, int q, int k, int32_t *ids, float *dist2, float *ret,
                                      const ListArgs &)
  {
    for (int i = 0; i < k; i++) {
      ids[(size_t)q * k + i] = (int32_t)(uint32_t)s.r.entry[i];
      dist2[(size_t)q * k + i] = __uint_as_float((uint32_t)(s.r.entry[i] >> 32));
    }
    if (ret) ret[q] = s.r.maxRadius2();
  }
};

// HeapCandidateList with a run-time k: the reference's flexible heap list
// over per-lane storage
template <> struct ResultOps<cukd::FlexHeapCandidateList> {
  struct State {
    uint64_t mem[64];
    cukd::FlexHeapCandidateList r;
    __device__ State() : r(mem, 1, -1.f) {}
  };
  __device__ static inline cukd::FlexHeapCandidateList &init(State &s, float cutoff, int k, int,
                                                             const ListArgs &)
  {
    new (&s.r) cukd::FlexHeapCandidateList(s.mem, k, cutoff);
    return s.r;
  }
  __device__ static inline void write(const State &s, int q, int k, int32_t *ids, float *dist2, float *ret,
                                      const ListArgs &)
  {
    for (int i = 0; i < k; i++) {
      ids[(size_t)q * k + i] = (int32_t)(uint32_t)s.mem[i];
      dist2[(size_t)q * k + i] = __uint_as_float((uint32_t)(s.mem[i] >> 32));
    }
    if (ret) ret[q] = s.r.maxRadius2();
  }
};

// Queries over a forest (many trees back to back in one array): query qi
// belongs to the segment s with seg_off[s] <= qi < seg_off[s+1] and runs
// against tree seg_tree[s] = elements tree_off[t] .. tree_off[t+1].
// seg_off == nullptr: one tree, the TreeView as given.
struct ForestQ {
  const int32_t *seg_off;
  const int32_t *seg_tree;
  int nseg;
  const int32_t *tree_off;
};

__device__ inline int segment_of(const int32_t *seg_off, int nseg, int qi)
{
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (seg_off[mid] <= qi) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// batch kernel: one query per lane
// ---------------------------------------------------------------------------
template <int D, int METHOD, class R>
__global__ __launch_bounds__(256) void k_query(TreeView t, const float *queries, int nq,
                                               const float *world_bounds, float cutoff,
                                               float eps, int k, const int32_t *order,
                                               int32_t *ids, float *dist2, float *ret,
                                               ListArgs la, ForestQ fq,
                                               unsigned long long *stats)
{
  // Workgroups are dispatched round-robin over the 8 XCDs, each with its
  // own L2: give every XCD one contiguous run of (Morton-ordered) queries,
  // so neighbouring queries share an L2 instead of fetching the same tree
  // nodes into eight of them.  A bijection of [0, gridDim.x).
  const int nb = (int)gridDim.x;
  const int xcd = (int)blockIdx.x & 7, q8 = nb >> 3, r8 = nb & 7;
  const int lb = xcd * q8 + (xcd < r8 ? xcd : r8) + ((int)blockIdx.x >> 3);
  const int i = lb * (int)blockDim.x + (int)threadIdx.x;
  if (i >= nq) return;
  const int qi = order ? order[i] : i;
  if (fq.seg_off) {
    const int tr = fq.seg_tree[segment_of(fq.seg_off, fq.nseg, qi)];
    const int b = fq.tree_off[tr];
    t.base += (size_t)b * t.stride;
    t.n = fq.tree_off[tr + 1] - b;
    if (world_bounds) world_bounds += (size_t)tr * 2 * D;
  }
  float q[D];
#pragma unroll
  for (int d = 0; d < D; d++) q[d] = queries[(size_t)qi * D + d];
  typename ResultOps<R>::State st;
  R &res = ResultOps<R>::init(st, cutoff, k, qi, la);
  unsigned visits = 0;
  TreeViewD<D> tv;
  static_cast<TreeView &>(tv) = t;
  if (METHOD == CUKD_STACK_BASED) {
    cukd::detail::traverse_stack_based<D>(res, q, tv, visits);
  } else if (METHOD == CUKD_STACK_FREE) {
    cukd::detail::traverse_stack_free<D>(res, q, tv, eps, visits);
  } else {
    float lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; d++) { lo[d] = world_bounds[d]; hi[d] = world_bounds[D + d]; }
    cukd::detail::traverse_cct<D>(res, q, lo, hi, tv, visits);
  }
  ResultOps<R>::write(st, qi, k, ids, dist2, ret, la);
  // CUKD_STATS (common.h:348-360): one count per visited node
  if (stats) atomicAdd(stats, (unsigned long long)visits);
}

// ---------------------------------------------------------------------------
// per-dimension launchers
// ---------------------------------------------------------------------------
struct QueryArgs {
  TreeView t;
  const float *queries;
  int nq;
  const float *world_bounds;
  float cutoff;
  float eps;
  int k;
  int list_kind;   // -1 = FCP, CUKD_LIST_FIXED, CUKD_LIST_HEAP or LIST_ACCUM
  const int32_t *order;
  int32_t *ids;
  float *dist2;
  float *ret;
  ListArgs la;
  ForestQ fq;
  unsigned long long *stats;
  int block = 128;  // workgroup size: 64, 128 or 256
};

// persistent multi-round lists (cukd_knn_accumulate)
enum { LIST_ACCUM = 100 };

template <int D, int METHOD, class R>
inline hipError_t launch_q(const QueryArgs &a, hipStream_t s)
{
  const int nt = a.block;
  const int blocks = (a.nq + nt - 1) / nt;
  hipLaunchKernelGGL((k_query<D, METHOD, R>), dim3(blocks), dim3(nt), 0, s, a.t, a.queries,
                     a.nq, a.world_bounds, a.cutoff, a.eps, a.k, a.order, a.ids, a.dist2,
                     a.ret, a.la, a.fq, a.stats);
  return hipGetLastError();
}

template <int D, int METHOD>
inline hipError_t launch_method(const QueryArgs &a, hipStream_t s)
{
  if (a.list_kind < 0) return launch_q<D, METHOD, cukd::FCPResult>(a, s);
  if (a.list_kind == CUKD_LIST_HEAP) return launch_q<D, METHOD, cukd::FlexHeapCandidateList>(a, s);
  if (a.list_kind == LIST_ACCUM) return launch_q<D, METHOD, AccumList>(a, s);
  switch (a.k) {
    case 1: return launch_q<D, METHOD, cukd::FixedCandidateList<1>>(a, s);
    case 4: return launch_q<D, METHOD, cukd::FixedCandidateList<4>>(a, s);
    case 8: return launch_q<D, METHOD, cukd::FixedCandidateList<8>>(a, s);
    case 16: return launch_q<D, METHOD, cukd::FixedCandidateList<16>>(a, s);
    default: return launch_q<D, METHOD, DynFixedList>(a, s);
  }
}

template <int D>
inline hipError_t launch_query_d(int method, const QueryArgs &a, hipStream_t s)
{
  if (method == CUKD_STACK_BASED) return launch_method<D, CUKD_STACK_BASED>(a, s);
  if (method == CUKD_STACK_FREE) return launch_method<D, CUKD_STACK_FREE>(a, s);
  return launch_method<D, CUKD_CCT>(a, s);
}

hipError_t launch_query(int D, int method, const QueryArgs &a, hipStream_t s);
hipError_t launch_knn_extract(const uint64_t *lists, int nq, int k, int32_t *ids, float *dist2,
                              float *radius, hipStream_t s);

// spatial-tree queries (spatial.hip)
struct SpatialBounds {
  float v[2 * CUKD_MAX_DIMS];
};
struct SpatialArgs {
  const cukd_spatial_node *nodes;
  const uint32_t *prim_ids;
  TreeView data;
  SpatialBounds bounds;
  const float *queries;
  int nq;
  float cutoff;
  int k;
  int list_kind;  // -1 = FCP
  int32_t *ids;
  float *dist2;
  float *ret;
};
hipError_t launch_spatial_query(int D, int method, const SpatialArgs &a, hipStream_t s);
int spatial_build_impl(cukd_spatial_tree *tree, const void *d_points, int n, const cukd_layout &L,
                       int leaf_threshold, hipStream_t s, const MemResource &mem);

}  // namespace cukd_amd
