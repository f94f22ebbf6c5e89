// query_impl.h -- batched FCP / kNN kernels over a left-balanced k-d tree.
//
// One query per lane.  The traversals and result types are the public
// device-inline ones of include/cukd/ (traverse.h, fcp.h, knn.h), run
// through an accessor over the run-time element layout, so the batched
// path and user kernels share one implementation; ids and dist2 bits
// equal the CPU oracle's, ties included.
#pragma once
#pragma clang fp contract(off)

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

// The run-time-k lists (k <= 64) keep their entries in LDS, in columns:
// entry i of the lane's query at col[i * stride], col = base + lane and
// stride = the workgroup size, so the lanes of a wave touch consecutive
// words.  (A per-lane array indexed at run time lives in scratch memory:
// k = 64 took 304 ms for 1M queries over a 10M-point tree.)  The largest
// kept key is also held in a register, so rejected candidates -- most of them
// -- touch no LDS.
struct LdsCol {
  uint64_t *p;
  int stride;
  __device__ inline uint64_t &operator[](int i) const { return p[(size_t)i * stride]; }
};

__device__ inline uint64_t list_key(float d2, uint32_t id)
{
  return ((uint64_t)__float_as_uint(d2) << 32) | id;
}

// the reference's max-heap push (knn.h:456-483, as detail::heap_push) over
// an LDS column; top = e[0]
__device__ inline void lds_heap_push(const LdsCol &e, int k, uint64_t v, uint64_t &top)
{
  if (v >= top) return;
  int pos = 0;
  for (;;) {
    uint64_t largest_val = ~0ull;
    const int c0 = 2 * pos + 1;
    int largest = k;
    if (c0 < k) {
      largest = c0;
      largest_val = e[c0];
    }
    const int c1 = c0 + 1;
    if (c1 < k) {
      const uint64_t v1 = e[c1];
      if (v1 > largest_val) {
        largest = c1;
        largest_val = v1;
      }
    }
    if (largest == k || largest_val < v) {
      e[pos] = v;
      break;
    }
    e[pos] = largest_val;
    pos = largest;
  }
  top = e[0];
}

// in-place heapsort of a max-heap column into ascending order
__device__ inline void lds_heap_sort(const LdsCol &e, int k)
{
  for (int end = k - 1; end > 0; end--) {
    const uint64_t v = e[end];
    e[end] = e[0];
    int pos = 0;
    for (;;) {
      int c = 2 * pos + 1;
      if (c >= end) break;
      uint64_t cv = e[c];
      if (c + 1 < end) {
        const uint64_t v1 = e[c + 1];
        if (v1 > cv) {
          c++;
          cv = v1;
        }
      }
      if (!(cv > v)) break;
      e[pos] = cv;
      pos = c;
    }
    e[pos] = v;
  }
}

// Run-time-k candidate lists over an LDS column, as a max-heap (top = the
// k-th smallest key).
//  SORTED = false: HeapCandidateList / FlexHeapCandidateList (knn.h:240-
//    287), written out in heap order as the reference leaves it.
//  SORTED = true: FixedCandidateList with a run-time k (knn.h:214-232).  The
//    fixed list keeps the k smallest keys offered so far, ascending, and
//    culls with the largest of them; a max-heap given the same candidates
//    holds the same k keys with the same largest one, so every cull and
//    every accepted candidate is the same, and one heapsort at the end
//    gives the fixed list's ascending order.  (A sorted insert moves up to
//    k - 1 words per accepted candidate; the heap moves log2(k).)
template <bool SORTED> struct LdsList {
  static constexpr bool kSorted = SORTED;
  LdsCol e;
  int k;
  uint64_t top;  // e[0]
  __device__ inline void init_top(uint64_t v) { top = v; }
  __device__ inline float maxRadius2() const { return __uint_as_float((uint32_t)(top >> 32)); }
  __device__ inline float initialCullDist2() const { return maxRadius2(); }
  __device__ inline float processCandidate(int id, float d2)
  {
    lds_heap_push(e, k, list_key(d2, (uint32_t)id), top);
    return maxRadius2();
  }
};
using DynFixedList = LdsList<true>;
using LdsHeapList = LdsList<false>;

// persistent per-query candidate lists in HBM (multi-round kNN)
struct ListArgs {
  uint64_t *lists;  // nq * k entries, query q at lists[q*k]
  int init;         // fill with (cutoff^2, -1) before the traversal
  int id_offset;    // added to every candidate id
};

// The reference's global-memory FlexHeapCandidateList of a multi-round
// query (samples/mpiHugeQuery.cu:75-88, knn.h:360-421): the k entries are
// loaded from HBM into the LDS column, updated with the same heap push,
// and stored back, so they end as the reference's in-place list would.
struct AccumList {
  LdsCol e;
  int k;
  int id_offset;
  uint64_t top;
  __device__ inline float maxRadius2() const { return __uint_as_float((uint32_t)(top >> 32)); }
  __device__ inline float initialCullDist2() const { return maxRadius2(); }
  __device__ inline float processCandidate(int id, float d2)
  {
    lds_heap_push(e, k, list_key(d2, (uint32_t)(id + id_offset)), top);
    return maxRadius2();
  }
};

// ---------------------------------------------------------------------------
// per-result-type set-up and output
// ---------------------------------------------------------------------------
// kLds: the type keeps k 64-bit words per query in the kernel's dynamic
// LDS (init gets the lane's column; the others ignore it)
template <class R> struct ResultOps;

template <> struct ResultOps<AccumList> {
  static constexpr bool kLds = true;
  struct State { AccumList r; };
  __device__ static inline AccumList &init(State &s, float cutoff, int k, int q, const ListArgs &la,
                                           LdsCol col)
  {
    s.r.e = col;
    s.r.k = k;
    s.r.id_offset = la.id_offset;
    const uint64_t *src = la.lists + (size_t)q * k;
    const uint64_t v = list_key(cutoff * cutoff, 0xffffffffu);
    for (int i = 0; i < k; i++) col[i] = la.init ? v : src[i];
    s.r.top = col[0];
    return s.r;
  }
  __device__ static inline void write(const State &s, int q, int k, int32_t *, float *, float *,
                                      const ListArgs &la)
  {
    uint64_t *dst = la.lists + (size_t)q * k;
    for (int i = 0; i < k; i++) dst[i] = s.r.e[i];
  }
};

template <> struct ResultOps<cukd::FCPResult> {
  static constexpr bool kLds = false;
  struct State { cukd::FCPResult r; };
  __device__ static inline cukd::FCPResult &init(State &s, float cutoff, int, int, const ListArgs &,
                                                 LdsCol)
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
  static constexpr bool kLds = false;
  struct State { cukd::FixedCandidateList<K> r; __device__ State() : r(0.f) {} };
  __device__ static inline cukd::FixedCandidateList<K> &init(State &s, float cutoff, int, int,
                                                             const ListArgs &, LdsCol)
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

// the run-time-k fixed and heap lists: (cutoff^2, -1) in every entry, and
// the entries written out in storage order
template <class L> struct LdsListOps {
  static constexpr bool kLds = true;
  struct State { L r; };
  __device__ static inline L &init(State &s, float cutoff, int k, int, const ListArgs &, LdsCol col)
  {
    s.r.e = col;
    s.r.k = k;
    const uint64_t v = list_key(cutoff * cutoff, 0xffffffffu);
    for (int i = 0; i < k; i++) col[i] = v;
    s.r.init_top(v);
    return s.r;
  }
  __device__ static inline void write(const State &s, int q, int k, int32_t *ids, float *dist2, float *ret,
                                      const ListArgs &)
  {
    if (L::kSorted) lds_heap_sort(s.r.e, k);
    for (int i = 0; i < k; i++) {
      const uint64_t x = s.r.e[i];
      ids[(size_t)q * k + i] = (int32_t)(uint32_t)x;
      dist2[(size_t)q * k + i] = __uint_as_float((uint32_t)(x >> 32));
    }
    if (ret) ret[q] = s.r.maxRadius2();
  }
};
template <bool SORTED> struct ResultOps<LdsList<SORTED>> : LdsListOps<LdsList<SORTED>> {};

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
  HIP_DYNAMIC_SHARED(uint64_t, s_lists);
  typename ResultOps<R>::State st;
  R &res = ResultOps<R>::init(st, cutoff, k, qi, la, LdsCol{s_lists + threadIdx.x, (int)blockDim.x});
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
// batch kernel: two stack-based queries per lane
// ---------------------------------------------------------------------------
// The stack-based traversal is a chain of dependent node loads, and a
// wave holds at most one load per lane in flight.  Here each lane steps
// two queries together: every iteration issues the node loads of both
// before using either, so twice as many loads are in flight per wave slot
// for little more than one query's registers.  Each query's steps are
// those of traverse_stack_based (include/cukd/traverse.h), one node load
// per step:
//   descend  visit node curr, mark its far child pending if the plane is
//            within the cull distance, go to the close child; past the
//            last node, pop
//   pop      re-read the parent of the deepest pending far child and enter
//            that child if its plane is within the cull distance, else pop
//            the next one
// so every query visits the same nodes in the same order with the same
// candidate updates, and its result is bit-identical to one query per lane.
//
// The step is written for few vector instructions, since a wave executes
// every path some lane takes: node byte offsets are 32-bit (trees under
// 4 GiB) and follow the descent by doubling, the level is counted rather
// than recomputed, and descend and pop share the plane test.
enum { SB_DESCEND = 0, SB_POP = 1, SB_DONE = 2 };

// round-robin split dim of level l (0 <= l < 32)
template <int D>
__device__ inline int rr_dim(int l)
{
  if ((D & (D - 1)) == 0) return l & (D - 1);
  if (D == 3) return l - (int)__umul24((unsigned)__umul24((unsigned)l, 11u) >> 5, 3u);  // l / 3 for l < 32
  return l % D;
}

struct SbSlot {
  int mode;
  int curr;          // descend: node to visit; pop: end of the path (>= n)
  int lvl;           // level of curr
  int pd;            // pop: depth of the far child being tested
  unsigned off;      // byte offset of the node loaded next (curr, or the far child's parent)
  unsigned pending;  // bit d: the far child at depth d of the path is still to be visited
  float cull;
};

template <int D, class R>
__device__ inline void sb_step(R &res, const float (&q)[D], const float (&p)[D], int n,
                               unsigned stride, SbSlot &s, unsigned &visits)
{
  if (s.mode == SB_DONE) return;
  const bool desc = s.mode == SB_DESCEND;
  const int dim = rr_dim<D>(desc ? s.lvl : s.pd - 1);
  const float nc = cukd::detail::pick<D>(p, dim);
  const float qc = cukd::detail::pick<D>(q, dim);
  if (desc) {
    visits++;
    s.cull = res.processCandidate(s.curr, cukd::detail::sqr_dist<D>(p, q));
  }
  const float diff = qc - nc;
  const bool near = diff * diff < s.cull;
  bool pop;
  if (desc) {
    const bool left_close = qc < nc;
    const int l = 2 * s.curr + 1;
    const int far = left_close ? l + 1 : l;
    if (near && far < n) s.pending |= 2u << s.lvl;
    s.curr = left_close ? l : l + 1;
    s.off = 2u * s.off + (left_close ? stride : 2u * stride);
    s.lvl++;
    pop = s.curr >= n;
  } else {
    pop = !near;
    if (near) {
      // the far child: sibling of the path's node at depth pd
      const int f = (((s.curr + 1) >> (s.lvl - s.pd)) ^ 1) - 1;
      s.off = 2u * s.off + ((f & 1) ? stride : 2u * stride);
      s.curr = f;
      s.lvl = s.pd;
      s.mode = SB_DESCEND;
    }
  }
  if (pop) {
    if (!s.pending) {
      s.mode = SB_DONE;
      return;
    }
    const int d = 31 - __clz((int)s.pending);
    s.pending &= ~(1u << d);
    s.pd = d;
    s.mode = SB_POP;
    s.off = (unsigned)(((s.curr + 1) >> (s.lvl - d + 1)) - 1) * stride;
  }
}

// lists for which large sorted batches run with fewer resident waves (FCP
// and k = 1 touch fewer nodes per query and measured faster at full
// occupancy)
template <class R> struct FewWaves { enum { ok = 0 }; };
template <int K> struct FewWaves<cukd::FixedCandidateList<K>> { enum { ok = K >= 4 }; };

// result types whose per-lane state is registers only (trees with
// round-robin split dims only)
template <class R> struct DualLane { enum { ok = 0 }; };
template <> struct DualLane<cukd::FCPResult> { enum { ok = 1 }; };
template <int K> struct DualLane<cukd::FixedCandidateList<K>> { enum { ok = 1 }; };

template <int D, class R, bool FOREST>
__global__ __launch_bounds__(256) void k_query2(TreeView t, const float *queries, int nq,
                                                float cutoff, int k, const int32_t *order,
                                                int32_t *ids, float *dist2, float *ret,
                                                ListArgs la, ForestQ fq,
                                                unsigned long long *stats)
{
  // XCD-aware workgroup order as in k_query; workgroup lb covers queries
  // [lb*2*nt, (lb+1)*2*nt), wave w of it the 128 from lb*2*nt + 128*w, and
  // lane l of the wave the two at +l and +64+l
  const int nb = (int)gridDim.x;
  const int xcd = (int)blockIdx.x & 7, q8 = nb >> 3, r8 = nb & 7;
  const int lb = xcd * q8 + (xcd < r8 ? xcd : r8) + ((int)blockIdx.x >> 3);
  const int i0 = lb * 2 * (int)blockDim.x + ((int)threadIdx.x >> 6) * 128 + ((int)threadIdx.x & 63);
  if (i0 >= nq) return;
  const bool has_b = i0 + 64 < nq;
  const int qa = order ? order[i0] : i0;
  const int qb = has_b ? (order ? order[i0 + 64] : i0 + 64) : qa;
  const unsigned stride = (unsigned)t.stride;
  const uint8_t *base = t.base + t.point_offset;
  unsigned tb_a = 0, tb_b = 0;  // byte offset of the slot's tree in the forest
  int n_a = t.n, n_b = t.n;
  if constexpr (FOREST) {
    const int ta = fq.seg_tree[segment_of(fq.seg_off, fq.nseg, qa)];
    const int tb = fq.seg_tree[segment_of(fq.seg_off, fq.nseg, qb)];
    const int oa = fq.tree_off[ta], ob = fq.tree_off[tb];
    tb_a = (unsigned)oa * stride;
    tb_b = (unsigned)ob * stride;
    n_a = fq.tree_off[ta + 1] - oa;
    n_b = fq.tree_off[tb + 1] - ob;
  }
  float pqa[D], pqb[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    pqa[d] = queries[(size_t)qa * D + d];
    pqb[d] = queries[(size_t)qb * D + d];
  }
  typename ResultOps<R>::State sta, stb;
  R &ra = ResultOps<R>::init(sta, cutoff, k, qa, la, LdsCol{});
  R &rb = ResultOps<R>::init(stb, cutoff, k, qb, la, LdsCol{});
  SbSlot sa{n_a > 0 ? SB_DESCEND : SB_DONE, 0, 0, 0, 0u, 0u, ra.initialCullDist2()};
  SbSlot sb{(has_b && n_b > 0) ? SB_DESCEND : SB_DONE, 0, 0, 0, 0u, 0u, rb.initialCullDist2()};
  unsigned visits = 0;
  while (sa.mode != SB_DONE || sb.mode != SB_DONE) {
    // a finished query re-reads the other's node, so both loads stay in
    // bounds and unconditional
    unsigned oa = tb_a + sa.off, ob = tb_b + sb.off;
    if (sa.mode == SB_DONE) oa = ob;
    if (sb.mode == SB_DONE) ob = oa;
    const float *src_a = (const float *)(base + oa);
    const float *src_b = (const float *)(base + ob);
    float pa[D], pb[D];
#pragma unroll
    for (int d = 0; d < D; d++) pa[d] = src_a[d];
#pragma unroll
    for (int d = 0; d < D; d++) pb[d] = src_b[d];
    sb_step<D>(ra, pqa, pa, n_a, stride, sa, visits);
    sb_step<D>(rb, pqb, pb, n_b, stride, sb, visits);
  }
  ResultOps<R>::write(sta, qa, k, ids, dist2, ret, la);
  if (has_b) ResultOps<R>::write(stb, qb, k, ids, dist2, ret, la);
  if (stats) atomicAdd(stats, (unsigned long long)visits);
}

// ---------------------------------------------------------------------------
// two queries per lane, pending plane distances kept in LDS
// ---------------------------------------------------------------------------
// The traversal above re-reads a far child's parent on pop to recompute its
// plane distance, which costs a loop iteration of its own. Here the plane
// distance is stored when the far child is pushed (one LDS word per depth,
// per query), so a pop is decided before the iteration's loads: a pop
// that passes the cull test descends into the far child in the same
// iteration, and only a pop that fails takes one. The visits, candidate
// updates and visit counts are exactly those of the traversal above
// (traverse-default-stack-based.h:58-107): the stored value is the same
// diff*diff of the same node and query, tested against the same current
// cull.
struct SdSlot {
  int mode;          // SB_DESCEND, SB_POP or SB_DONE
  int curr;          // descend: node to visit; pop: end of the path (>= n)
  int lvl;           // level of curr
  unsigned off;      // descend: byte offset of curr
  unsigned pending;  // bit d: the far child at depth d of the path is still to be visited
  float topd;        // plane distance of the deepest pending far child
  float cull;
};

// rows of per-query LDS words the traversal needs: a far child at depth d
// exists only if 2^d <= n, and depth d is stored in row d - 1
__host__ __device__ inline int sd_rows(int n)
{
  int r = 0;
  while (r < 31 && (2ll << r) <= (long long)n) r++;
  return r > 0 ? r : 1;
}

// pop the deepest pending far child: descend into it if its stored plane
// distance is still below the cull, and reload the next stored distance
__device__ inline void sd_pop(SdSlot &s, const float *col, int row_stride, unsigned stride)
{
  if (s.mode != SB_POP) return;
  const int d = 31 - __clz((int)s.pending);
  s.pending &= ~(1u << d);
  if (s.topd < s.cull) {
    const int f = (((s.curr + 1) >> (s.lvl - d)) ^ 1) - 1;
    s.curr = f;
    s.lvl = d;
    s.off = (unsigned)f * stride;
    s.mode = SB_DESCEND;
  } else if (!s.pending) {
    s.mode = SB_DONE;
  }
  if (s.pending) s.topd = col[(30 - __clz((int)s.pending)) * row_stride];
}

// visit curr, push its far child (with its plane distance) if it may hold
// closer points, and step to the close child
template <int D, class R>
__device__ inline void sd_visit(R &res, const float (&q)[D], const float (&p)[D], int n,
                                unsigned stride, SdSlot &s, float *col, int row_stride,
                                unsigned &visits)
{
  if (s.mode != SB_DESCEND) return;
  visits++;
  s.cull = res.processCandidate(s.curr, cukd::detail::sqr_dist<D>(p, q));
  const int dim = rr_dim<D>(s.lvl);
  const float nc = cukd::detail::pick<D>(p, dim);
  const float qc = cukd::detail::pick<D>(q, dim);
  const float diff = qc - nc;
  const float dd = diff * diff;
  const bool left_close = qc < nc;
  const int l = 2 * s.curr + 1;
  const int far = left_close ? l + 1 : l;
  if (dd < s.cull && far < n) {
    s.pending |= 2u << s.lvl;
    col[s.lvl * row_stride] = dd;
    s.topd = dd;
  }
  s.curr = left_close ? l : l + 1;
  s.off = 2u * s.off + (left_close ? stride : 2u * stride);
  s.lvl++;
  if (s.curr >= n) s.mode = s.pending ? SB_POP : SB_DONE;
}

// Dynamic LDS: sd_rows(n) rows of 2 * blockDim.x floats (row r holds depth
// r + 1 of every slot of the workgroup; each lane only touches its own two
// columns, so no barrier is needed).
template <int D, class R, bool FOREST>
__global__ __launch_bounds__(256) void k_query2s(TreeView t, const float *queries, int nq,
                                                 float cutoff, int k, const int32_t *order,
                                                 int32_t *ids, float *dist2, float *ret,
                                                 ListArgs la, ForestQ fq,
                                                 unsigned long long *stats)
{
  HIP_DYNAMIC_SHARED(float, s_pd);
  const int nb = (int)gridDim.x;
  const int xcd = (int)blockIdx.x & 7, q8 = nb >> 3, r8 = nb & 7;
  const int lb = xcd * q8 + (xcd < r8 ? xcd : r8) + ((int)blockIdx.x >> 3);
  const int i0 = lb * 2 * (int)blockDim.x + ((int)threadIdx.x >> 6) * 128 + ((int)threadIdx.x & 63);
  if (i0 >= nq) return;
  const bool has_b = i0 + 64 < nq;
  const int qa = order ? order[i0] : i0;
  const int qb = has_b ? (order ? order[i0 + 64] : i0 + 64) : qa;
  const unsigned stride = (unsigned)t.stride;
  const uint8_t *base = t.base + t.point_offset;
  unsigned tb_a = 0, tb_b = 0;  // byte offset of the slot's tree in the forest
  int n_a = t.n, n_b = t.n;
  if constexpr (FOREST) {
    const int ta = fq.seg_tree[segment_of(fq.seg_off, fq.nseg, qa)];
    const int tb = fq.seg_tree[segment_of(fq.seg_off, fq.nseg, qb)];
    const int oa = fq.tree_off[ta], ob = fq.tree_off[tb];
    tb_a = (unsigned)oa * stride;
    tb_b = (unsigned)ob * stride;
    n_a = fq.tree_off[ta + 1] - oa;
    n_b = fq.tree_off[tb + 1] - ob;
  }
  const int row_stride = 2 * (int)blockDim.x;
  float *col_a = s_pd + threadIdx.x;
  float *col_b = s_pd + blockDim.x + threadIdx.x;
  float pqa[D], pqb[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    pqa[d] = queries[(size_t)qa * D + d];
    pqb[d] = queries[(size_t)qb * D + d];
  }
  typename ResultOps<R>::State sta, stb;
  R &ra = ResultOps<R>::init(sta, cutoff, k, qa, la, LdsCol{});
  R &rb = ResultOps<R>::init(stb, cutoff, k, qb, la, LdsCol{});
  SdSlot sa{n_a > 0 ? SB_DESCEND : SB_DONE, 0, 0, 0u, 0u, 0.f, ra.initialCullDist2()};
  SdSlot sb{(has_b && n_b > 0) ? SB_DESCEND : SB_DONE, 0, 0, 0u, 0u, 0.f, rb.initialCullDist2()};
  unsigned visits = 0;
  while (sa.mode != SB_DONE || sb.mode != SB_DONE) {
    sd_pop(sa, col_a, row_stride, stride);
    sd_pop(sb, col_b, row_stride, stride);
    // a slot that does not descend this iteration re-reads the other's
    // node (or the root), so both loads stay in bounds and unconditional
    const bool da = sa.mode == SB_DESCEND, db = sb.mode == SB_DESCEND;
    const unsigned oa = da ? tb_a + sa.off : (db ? tb_b + sb.off : 0u);
    const unsigned ob = db ? tb_b + sb.off : oa;
    const float *src_a = (const float *)(base + oa);
    const float *src_b = (const float *)(base + ob);
    float pa[D], pb[D];
#pragma unroll
    for (int d = 0; d < D; d++) pa[d] = src_a[d];
#pragma unroll
    for (int d = 0; d < D; d++) pb[d] = src_b[d];
    sd_visit<D>(ra, pqa, pa, n_a, stride, sa, col_a, row_stride, visits);
    sd_visit<D>(rb, pqb, pb, n_b, stride, sb, col_b, row_stride, visits);
  }
  ResultOps<R>::write(sta, qa, k, ids, dist2, ret, la);
  if (has_b) ResultOps<R>::write(stb, qb, k, ids, dist2, ret, la);
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
  int lanes = 1;    // queries per lane: 1, or 2 (stack-based, register lists)
  uint64_t tree_bytes = 0;  // bytes of the tree array (forests: all trees)
  int max_n = 0;            // forests: points of the largest tree (LDS rows of k_query2s)
  bool presorted = false;   // queries already in Morton order (CUKD_QUERY_PRESORTED)
};

// persistent multi-round lists (cukd_knn_accumulate)
enum { LIST_ACCUM = 100 };

// gfx950 LDS per CU, and the batch size from which the two-query kernel
// runs with fewer resident waves (a build option so that small test
// batches can take that path too)
#ifndef CUKD_QUERY2_FEW_WAVES_MIN_NQ
#define CUKD_QUERY2_FEW_WAVES_MIN_NQ (1 << 19)
#endif
enum { LDS_PER_CU = 163840, QUERY2_FEW_WAVES_MIN_NQ = CUKD_QUERY2_FEW_WAVES_MIN_NQ };

template <int D, int METHOD, class R>
inline hipError_t launch_q(const QueryArgs &a, hipStream_t s)
{
  const int nt = a.block;
  if constexpr (METHOD == CUKD_STACK_BASED && DualLane<R>::ok) {
    if (a.lanes == 2 && a.t.dim_offset < 0) {
      const int blocks = (a.nq + 2 * nt - 1) / (2 * nt);
      // Large Morton-ordered batches run with fewer resident waves:
      // workgroups are handed out in order (XCD-aware, along the Morton
      // order), so fewer waves in flight means each XCD's active queries
      // cover a smaller region and their nodes stay in its L2.  About 14
      // waves per CU measured best for kNN-8 at 1M queries (10 % faster than
      // 32, with 2.5x less L2-miss traffic); the dynamic LDS, unused by the
      // kernel, only limits how many workgroups share a CU.
      const bool few = FewWaves<R>::ok && (a.order || a.presorted) && nt == 128 &&
                       a.nq >= QUERY2_FEW_WAVES_MIN_NQ;
      const size_t lds = few ? LDS_PER_CU / 7 : 0;
#ifndef CUKD_NO_STORED_PLANES
      // Large batches, sorted or not, keep the pending plane distances in
      // LDS.  The LDS rows for the deepest tree also bound the resident
      // waves (6-7 workgroups per CU at 10M points), which suits unsorted
      // batches too: fewer queries in flight thrash the L2s less.  A
      // forest without a known largest tree keeps the kernel that re-reads
      // parents.
      if (nt == 128 && a.nq >= QUERY2_FEW_WAVES_MIN_NQ && (!a.fq.seg_off || a.max_n > 0)) {
        const size_t need =
            (size_t)sd_rows(a.fq.seg_off ? a.max_n : a.t.n) * 2 * nt * sizeof(float);
        const size_t shm = need > lds ? need : lds;
        if (a.fq.seg_off)
          hipLaunchKernelGGL((k_query2s<D, R, true>), dim3(blocks), dim3(nt), shm, s, a.t,
                             a.queries, a.nq, a.cutoff, a.k, a.order, a.ids, a.dist2, a.ret, a.la,
                             a.fq, a.stats);
        else
          hipLaunchKernelGGL((k_query2s<D, R, false>), dim3(blocks), dim3(nt), shm, s, a.t,
                             a.queries, a.nq, a.cutoff, a.k, a.order, a.ids, a.dist2, a.ret, a.la,
                             a.fq, a.stats);
        return hipGetLastError();
      }
#endif
      if (a.fq.seg_off)
        hipLaunchKernelGGL((k_query2<D, R, true>), dim3(blocks), dim3(nt), lds, s, a.t, a.queries,
                           a.nq, a.cutoff, a.k, a.order, a.ids, a.dist2, a.ret, a.la, a.fq, a.stats);
      else
        hipLaunchKernelGGL((k_query2<D, R, false>), dim3(blocks), dim3(nt), lds, s, a.t, a.queries,
                           a.nq, a.cutoff, a.k, a.order, a.ids, a.dist2, a.ret, a.la, a.fq, a.stats);
      return hipGetLastError();
    }
  }
  const int blocks = (a.nq + nt - 1) / nt;
  // run-time-k lists: k words per lane of dynamic LDS (k <= 64 and nt <=
  // 256 keep it within one CU's 160 KiB)
  const size_t shm = ResultOps<R>::kLds ? (size_t)a.k * sizeof(uint64_t) * nt : 0;
  if (shm > LDS_PER_CU) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_query<D, METHOD, R>), dim3(blocks), dim3(nt), shm, s, a.t, a.queries,
                     a.nq, a.world_bounds, a.cutoff, a.eps, a.k, a.order, a.ids, a.dist2,
                     a.ret, a.la, a.fq, a.stats);
  return hipGetLastError();
}

template <int D, int METHOD>
inline hipError_t launch_method(const QueryArgs &a, hipStream_t s)
{
  if (a.list_kind < 0) return launch_q<D, METHOD, cukd::FCPResult>(a, s);
  if (a.list_kind == CUKD_LIST_HEAP) return launch_q<D, METHOD, LdsHeapList>(a, s);
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
  unsigned long long *stats;  // node + leaf-point visits (fcp.h:294,327), nullable
};
hipError_t launch_spatial_query(int D, int method, const SpatialArgs &a, hipStream_t s);
int spatial_build_impl(cukd_spatial_tree *tree, const void *d_points, int n, const cukd_layout &L,
                       int leaf_threshold, hipStream_t s, const MemResource &mem);

}  // namespace cukd_amd
