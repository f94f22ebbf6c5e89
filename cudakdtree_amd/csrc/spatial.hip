// spatial.hip -- spatial k-d tree build and queries for gfx950.
//
// Build (spatial-kdtree.h:496-612 in the reference), one round per tree
// level over all open nodes at once:
//   decide   : leaf if count <= threshold, else split the widest centroid
//              extent at its centre (ties pick the earlier dim; a centre
//              equal to either end makes a degenerate "dim -1" split)
//   number   : children of this round's split nodes get consecutive ids in
//              ascending parent order (exclusive scan; the reference takes
//              atomicAdd order, so its numbering is not reproducible)
//   partition: each node's prims stay contiguous and in ascending prim id;
//              a stable partition sends them to the two children ("coord >=
//              pos" goes right; a degenerate split alternates by rank)
//   bounds   : children's centroid boxes grow with int-encoded float
//              atomics (order independent)
// then primIDs are laid out leaf by leaf in node-id order.  The host reads
// one counter per round to know whether another round is needed, as the
// reference does (spatial-kdtree.h:545-567).
//
// Queries: stack-based and closest-corner traversal with the reference's
// 50-entry stack and overflow behaviour (fcp.h:256-427, knn.h:566-752).
#pragma clang fp contract(off)
#include <new>

#include "cukd/spatial-kdtree.h"
#include "cukd_common.h"
#include "query_impl.h"

namespace cukd_amd {

// AtomicBox encoding (spatial-kdtree.h:172-188): float -> int32 whose
// signed order is the float order (-0.0 sorts below +0.0)
__device__ inline int32_t abox_enc(float f)
{
  int32_t b = __float_as_int(f);
  if (b & (int32_t)0x80000000) b ^= 0x7fffffff;
  return b;
}
__device__ inline float abox_dec(int32_t b)
{
  if (b & (int32_t)0x80000000) b ^= 0x7fffffff;
  return __int_as_float(b);
}

// A coordinate as the builder sees it: int coordinates go through the
// reference's implicit int -> float conversion (AtomicBox::encode(float),
// 'const float center = get_coord(point,split.dim)',
// spatial-kdtree.h:172-188, 427-429).
__device__ inline float sp_coord(const uint8_t *pts, uint32_t i, const cukd_layout &L, int d)
{
  const uint32_t b = load_coord_bits(pts + (size_t)i * L.stride, L, d);
  return L.scalar_type == CUKD_SCALAR_F32 ? __uint_as_float(b) : (float)(int32_t)b;
}

// Node::pos and the bounds hold scalar_t: a float result stored into an
// int one is truncated by C's conversion for int points
// (spatial-kdtree.h:75-77, 163-168, 489).
__device__ inline uint32_t sp_store(float f, bool isf)
{
  return isf ? __float_as_uint(f) : (uint32_t)(int32_t)f;
}

enum { SP_OPEN = 0, SP_LEAF = 1, SP_SPLIT = 2, SP_INNER = 3 };

struct SpNode {
  int32_t lo[CUKD_MAX_DIMS], hi[CUKD_MAX_DIMS];
  uint32_t count;
  uint32_t start;    // first position of this node's prims
  uint32_t offset;   // first child (inner)
  int32_t dim;
  float pos;
  int32_t state;
};

// ---------------------------------------------------------------------------
// exclusive scan of uint32 (multi-block, recursive over block sums)
// ---------------------------------------------------------------------------
enum { SCAN_NT = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_NT * SCAN_ITEMS };

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *s, uint32_t &total)
{
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < SCAN_NT; off <<= 1) {
    const uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  total = s[SCAN_NT - 1];
  const uint32_t r = s[threadIdx.x] - v;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_tiles(const uint32_t *in, uint32_t *out,
                                                        uint32_t *tile_sums, long long n)
{
  __shared__ uint32_t s[SCAN_NT];
  const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS], sum = 0;
  for (int i = 0; i < SCAN_ITEMS; i++) {
    v[i] = base + i < n ? in[base + i] : 0u;
    sum += v[i];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, s, total);
  for (int i = 0; i < SCAN_ITEMS; i++) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_add(uint32_t *out, const uint32_t *tile_offsets,
                                                      long long n)
{
  const long long base = (long long)blockIdx.x * SCAN_TILE;
  const uint32_t add = tile_offsets[blockIdx.x];
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_NT)
    if (base + i < n) out[base + i] += add;
}

// scratch: at least scan_scratch_words(n) uint32
size_t scan_scratch_words(long long n)
{
  size_t w = 0;
  while (n > 1) {
    n = (n + SCAN_TILE - 1) / SCAN_TILE;
    w += (size_t)n + 1;
  }
  return w + 1;
}

void scan_exclusive(const uint32_t *in, uint32_t *out, long long n, uint32_t *scratch,
                    hipStream_t s)
{
  if (n <= 0) return;
  const long long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  uint32_t *sums = scratch;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(SCAN_NT), 0, s, in, out, sums, n);
  if (tiles > 1) {
    scan_exclusive(sums, sums, tiles, scratch + tiles + 1, s);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(SCAN_NT), 0, s, out,
                       (const uint32_t *)sums, n);
  }
}

// ---------------------------------------------------------------------------
// build kernels
// ---------------------------------------------------------------------------
__global__ void k_sp_init(SpNode *nodes, uint32_t *pos_prim, uint32_t *pos_node, int n, int D)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 2) {
    SpNode &t = nodes[i];
    for (int d = 0; d < D; d++) { t.lo[d] = abox_enc(3.402823466e+38f); t.hi[d] = abox_enc(-3.402823466e+38f); }
    t.count = i == 0 ? (uint32_t)n : 0u;
    t.start = 0;
    t.offset = 0;
    t.dim = 0;
    t.pos = 0.f;
    t.state = i == 0 ? SP_OPEN : SP_LEAF;
  }
  if (i < n) {
    pos_prim[i] = (uint32_t)i;
    pos_node[i] = 0u;
  }
}

__global__ __launch_bounds__(256) void k_sp_root_bounds(SpNode *nodes, const uint8_t *pts, int n,
                                                        cukd_layout L)
{
  __shared__ int32_t s_lo[CUKD_MAX_DIMS][256], s_hi[CUKD_MAX_DIMS][256];
  const int D = L.num_dims;
  int32_t lo[CUKD_MAX_DIMS], hi[CUKD_MAX_DIMS];
  for (int d = 0; d < D; d++) { lo[d] = 0x7fffffff; hi[d] = (int32_t)0x80000000; }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    for (int d = 0; d < D; d++) {
      const int32_t e = abox_enc(sp_coord(pts, (uint32_t)i, L, d));
      lo[d] = min(lo[d], e);
      hi[d] = max(hi[d], e);
    }
  }
  for (int d = 0; d < D; d++) { s_lo[d][threadIdx.x] = lo[d]; s_hi[d][threadIdx.x] = hi[d]; }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int d = 0; d < D; d++) {
        s_lo[d][threadIdx.x] = min(s_lo[d][threadIdx.x], s_lo[d][threadIdx.x + off]);
        s_hi[d][threadIdx.x] = max(s_hi[d][threadIdx.x], s_hi[d][threadIdx.x + off]);
      }
    __syncthreads();
  }
  if ((int)threadIdx.x < D) {
    atomicMin(&nodes[0].lo[threadIdx.x], s_lo[threadIdx.x][0]);
    atomicMax(&nodes[0].hi[threadIdx.x], s_hi[threadIdx.x][0]);
  }
}

// selectSplits (spatial-kdtree.h:316-396) for the open nodes [a, b)
__global__ void k_sp_decide(SpNode *nodes, uint32_t *split_flag, int a, int b, int D,
                            uint32_t threshold)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (a + i >= b) return;
  SpNode &t = nodes[a + i];
  uint32_t is_split = 0;
  if (t.count <= threshold) {
    t.state = SP_LEAF;
    t.offset = 0xffffffffu;
    t.dim = 0;
    t.pos = 0.f;
  } else {
    float widest_w = 0.f;
    int widest = -1;
    for (int d = 0; d < D; d++) {
      const float w = abox_dec(t.hi[d]) - abox_dec(t.lo[d]);
      if (w <= widest_w) continue;
      widest_w = w;
      widest = d;
    }
    float pos = 0.f;
    if (widest >= 0) {
      const float lo = abox_dec(t.lo[widest]), hi = abox_dec(t.hi[widest]);
      pos = 0.5f * (lo + hi);
      if (pos == lo || pos == hi) widest = -1;
    }
    if (widest < 0) pos = abox_dec(t.lo[0]);
    t.dim = widest;
    t.pos = pos;
    t.state = SP_SPLIT;
    is_split = 1;
  }
  split_flag[i] = is_split;
}

// child ids and empty child boxes
__global__ void k_sp_alloc(SpNode *nodes, const uint32_t *split_excl, int a, int b, int next,
                           int D)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (a + i >= b) return;
  SpNode &t = nodes[a + i];
  if (t.state != SP_SPLIT) return;
  t.offset = (uint32_t)next + 2u * split_excl[i];
  for (int side = 0; side < 2; side++) {
    SpNode &c = nodes[t.offset + side];
    for (int d = 0; d < D; d++) { c.lo[d] = abox_enc(3.402823466e+38f); c.hi[d] = abox_enc(-3.402823466e+38f); }
    c.state = SP_OPEN;
    c.offset = 0;
    c.dim = 0;
    c.pos = 0.f;
    c.count = 0;
    c.start = t.start;
  }
}

// updatePrims (spatial-kdtree.h:398-436) part 1: side of every prim
__global__ void k_sp_classify(const SpNode *nodes, const uint32_t *pos_prim,
                              const uint32_t *pos_node, uint32_t *left_flag, const uint8_t *pts,
                              int n, cukd_layout L)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const SpNode &t = nodes[pos_node[p]];
  uint32_t left = 0;
  if (t.state == SP_SPLIT) {
    int side;
    if (t.dim < 0) {
      side = (int)(((uint32_t)p - t.start) & 1u);
    } else {
      side = sp_coord(pts, pos_prim[p], L, t.dim) >= t.pos;
    }
    left = side == 0;
  }
  left_flag[p] = left;
}

// updatePrims part 2: stable partition into the children, child boxes.
// The boxes grow with one atomic per (child, word) and wave: the lanes of a
// wave hold consecutive positions, so nearly always one or two children,
// and per-prim atomics on the few children of the top levels serialise
// (min and max are order-independent, so the boxes are the same).
__global__ __launch_bounds__(256) void k_sp_scatter(SpNode *nodes, const uint32_t *pos_prim,
                                                    const uint32_t *pos_node,
                                                    const uint32_t *left_flag,
                                                    const uint32_t *left_excl, uint32_t *npos_prim,
                                                    uint32_t *npos_node, const uint8_t *pts, int n,
                                                    cukd_layout L)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int D = L.num_dims;
  const int lane = (int)threadIdx.x & 63;
  uint32_t mychild = 0xffffffffu;  // none: past the end, or not split this round
  int32_t e[CUKD_MAX_DIMS];
#pragma unroll
  for (int d = 0; d < CUKD_MAX_DIMS; d++) e[d] = 0;
  if (p < n) {
    const uint32_t node = pos_node[p];
    const SpNode &t = nodes[node];
    const uint32_t prim = pos_prim[p];
    if (t.state != SP_SPLIT) {
      npos_prim[p] = prim;
      npos_node[p] = node;
    } else {
      const uint32_t end = t.start + t.count;
      const uint32_t left_end =
          end < (uint32_t)n ? left_excl[end] : left_excl[n - 1] + left_flag[n - 1];
      const uint32_t lcount = left_end - left_excl[t.start];
      const uint32_t lrank = left_excl[p] - left_excl[t.start];
      uint32_t np, child;
      if (left_flag[p]) {
        np = t.start + lrank;
        child = t.offset;
      } else {
        np = t.start + lcount + ((uint32_t)p - t.start - lrank);
        child = t.offset + 1;
      }
      npos_prim[np] = prim;
      npos_node[np] = child;
      SpNode &c = nodes[child];
      if (lrank == 0 && left_flag[p]) {
        c.count = lcount;
        c.start = t.start;
      }
      if (!left_flag[p] && (uint32_t)p - t.start - lrank == 0) {
        c.count = t.count - lcount;
        c.start = t.start + lcount;
      }
      mychild = child;
#pragma unroll
      for (int d = 0; d < CUKD_MAX_DIMS; d++)
        if (d < D) e[d] = abox_enc(sp_coord(pts, prim, L, d));
    }
  }
  // every lane of the wave takes part: one child per trip
  uint64_t active = __ballot(mychild != 0xffffffffu);
  while (active) {
    const int leader = __ffsll((unsigned long long)active) - 1;
    const uint32_t c = (uint32_t)__shfl((int)mychild, leader);
    const bool peer = mychild == c;
    active &= ~__ballot(peer);
#pragma unroll
    for (int d = 0; d < CUKD_MAX_DIMS; d++) {
      if (d >= D) break;
      int32_t lo = peer ? e[d] : 0x7fffffff, hi = peer ? e[d] : (int32_t)0x80000000;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o));
        hi = max(hi, __shfl_xor(hi, o));
      }
      if (lane == leader) {
        atomicMin(&nodes[c].lo[d], lo);
        atomicMax(&nodes[c].hi[d], hi);
      }
    }
  }
}

// split nodes of this round become inner nodes; count the next round
__global__ void k_sp_close(SpNode *nodes, const uint32_t *split_flag, const uint32_t *split_excl,
                           int a, int b, uint32_t *num_split)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (a + i >= b) return;
  SpNode &t = nodes[a + i];
  if (t.state == SP_SPLIT) t.state = SP_INNER;
  if (a + i == b - 1) *num_split = split_excl[i] + split_flag[i];
}

// leaf offsets in node-id order, then primIDs leaf by leaf
__global__ void k_sp_leaf_counts(const SpNode *nodes, uint32_t *cnt, int num_nodes)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_nodes) return;
  cnt[i] = nodes[i].state == SP_LEAF ? nodes[i].count : 0u;
}

__global__ void k_sp_write_prims(const SpNode *nodes, const uint32_t *leaf_off,
                                 const uint32_t *pos_prim, const uint32_t *pos_node,
                                 uint32_t *prim_ids, int n)
{
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t node = pos_node[p];
  prim_ids[leaf_off[node] + ((uint32_t)p - nodes[node].start)] = pos_prim[p];
}

// writeNodes (spatial-kdtree.h:477-494): dim -1 is stored as 0
__global__ void k_sp_write_nodes(const SpNode *nodes, const uint32_t *leaf_off,
                                 cukd_spatial_node *out, int num_nodes, bool isf)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= num_nodes) return;
  const SpNode &t = nodes[i];
  cukd_spatial_node o;
  if (t.state == SP_LEAF) {
    o.offset = t.count > 0 ? leaf_off[i] : (i == 1 ? 0u : 0xffffffffu);
    o.count = (uint16_t)t.count;
    o.dim = 0;
    o.pos = 0.f;
  } else {
    o.offset = t.offset;
    o.count = 0;
    o.dim = (int16_t)(t.dim < 0 ? 0 : t.dim);
    o.ipos = (int32_t)sp_store(t.pos, isf);
  }
  out[i] = o;
}

// saveBounds (spatial-kdtree.h:464-473): box_t of scalar_t
__global__ void k_sp_bounds_out(const SpNode *nodes, uint32_t *bounds, int D, bool isf)
{
  const int d = threadIdx.x;
  if (d < D) {
    bounds[d] = sp_store(abox_dec(nodes[0].lo[d]), isf);
    bounds[D + d] = sp_store(abox_dec(nodes[0].hi[d]), isf);
  }
}

int spatial_build_impl(cukd_spatial_tree *tree, const void *d_points, int n, const cukd_layout &L,
                       int leaf_threshold, hipStream_t s, const MemResource &mem)
{
  const int D = L.num_dims;
  const uint32_t thr = leaf_threshold <= 0 ? 8u : (uint32_t)leaf_threshold;
  const size_t max_nodes = 2 * (size_t)(n > 0 ? n : 1) + 2;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  const size_t scan_words = scan_scratch_words((long long)(max_nodes > nn ? max_nodes : nn));
  size_t o = 0;
  auto take = [&](size_t bytes) { const size_t at = o; o = (o + bytes + 255) & ~(size_t)255; return at; };
  const size_t o_nodes = take(max_nodes * sizeof(SpNode));
  const size_t o_pp0 = take(nn * 4), o_pn0 = take(nn * 4), o_pp1 = take(nn * 4), o_pn1 = take(nn * 4);
  const size_t o_nflag = take(max_nodes * 4), o_nexcl = take(max_nodes * 4);
  const size_t o_pflag = take(nn * 4), o_pexcl = take(nn * 4);
  const size_t o_scan = take(scan_words * 4), o_misc = take(256);
  uint8_t *tmp = nullptr;
  if (mem.alloc((void **)&tmp, o, s) != 0) return CUKD_ERR_ALLOC;
  SpNode *nodes = (SpNode *)(tmp + o_nodes);
  uint32_t *pp[2] = {(uint32_t *)(tmp + o_pp0), (uint32_t *)(tmp + o_pp1)};
  uint32_t *pn[2] = {(uint32_t *)(tmp + o_pn0), (uint32_t *)(tmp + o_pn1)};
  uint32_t *nflag = (uint32_t *)(tmp + o_nflag), *nexcl = (uint32_t *)(tmp + o_nexcl);
  uint32_t *pflag = (uint32_t *)(tmp + o_pflag), *pexcl = (uint32_t *)(tmp + o_pexcl);
  uint32_t *scratch = (uint32_t *)(tmp + o_scan);
  uint32_t *num_split = (uint32_t *)(tmp + o_misc);
  uint32_t *bounds_dev = (uint32_t *)(tmp + o_misc + 64);
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const uint8_t *pts = (const uint8_t *)d_points;
  int rc = CUKD_OK;
  hipError_t e = hipSuccess;

  hipLaunchKernelGGL(k_sp_init, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, nodes, pp[0],
                     pn[0], n, D);
  if (n > 0) {
    int blocks = (n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(k_sp_root_bounds, dim3(blocks), dim3(256), 0, s, nodes, pts, n, L);
  }
  // open nodes [a, b); node 1 is the reference's reserved empty leaf, so
  // the first children are numbered from 2
  int cur = 0, a = 0, b = 1, next = 2;
  const int gp = (n + 255) / 256;
  while (a < b) {
    const int cnt = b - a;
    const int g = (cnt + 255) / 256;
    hipLaunchKernelGGL(k_sp_decide, dim3(g), dim3(256), 0, s, nodes, nflag, a, b, D, thr);
    scan_exclusive(nflag, nexcl, cnt, scratch, s);
    hipLaunchKernelGGL(k_sp_alloc, dim3(g), dim3(256), 0, s, nodes, (const uint32_t *)nexcl, a, b, next,
                       D);
    if (n > 0) {
      hipLaunchKernelGGL(k_sp_classify, dim3(gp), dim3(256), 0, s, (const SpNode *)nodes,
                         (const uint32_t *)pp[cur], (const uint32_t *)pn[cur], pflag, pts, n, L);
      scan_exclusive(pflag, pexcl, n, scratch, s);
      hipLaunchKernelGGL(k_sp_scatter, dim3(gp), dim3(256), 0, s, nodes, (const uint32_t *)pp[cur],
                         (const uint32_t *)pn[cur], (const uint32_t *)pflag, (const uint32_t *)pexcl,
                         pp[cur ^ 1], pn[cur ^ 1], pts, n, L);
      cur ^= 1;
    }
    hipLaunchKernelGGL(k_sp_close, dim3(g), dim3(256), 0, s, nodes, (const uint32_t *)nflag,
                       (const uint32_t *)nexcl, a, b, num_split);
    uint32_t h_split = 0;
    e = hipMemcpyAsync(&h_split, num_split, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) break;
    a = next;
    b = next + 2 * (int)h_split;
    next = b;
  }
  const int num_nodes = next;
  cukd_spatial_node *out_nodes = nullptr;
  uint32_t *out_prims = nullptr;
  if (e == hipSuccess) {
    if (mem.alloc((void **)&out_nodes, (size_t)num_nodes * sizeof(cukd_spatial_node), s) != 0 ||
        mem.alloc((void **)&out_prims, nn * 4, s) != 0) {
      rc = CUKD_ERR_ALLOC;
    } else {
      const int g = (num_nodes + 255) / 256;
      hipLaunchKernelGGL(k_sp_leaf_counts, dim3(g), dim3(256), 0, s, (const SpNode *)nodes, nflag,
                         num_nodes);
      scan_exclusive(nflag, nexcl, num_nodes, scratch, s);
      if (n > 0)
        hipLaunchKernelGGL(k_sp_write_prims, dim3(gp), dim3(256), 0, s, (const SpNode *)nodes,
                           (const uint32_t *)nexcl, (const uint32_t *)pp[cur],
                           (const uint32_t *)pn[cur], out_prims, n);
      hipLaunchKernelGGL(k_sp_write_nodes, dim3(g), dim3(256), 0, s, (const SpNode *)nodes,
                         (const uint32_t *)nexcl, out_nodes, num_nodes, isf);
      hipLaunchKernelGGL(k_sp_bounds_out, dim3(1), dim3(64), 0, s, (const SpNode *)nodes,
                         bounds_dev, D, isf);
      uint32_t hb[2 * CUKD_MAX_DIMS];
      e = hipMemcpyAsync(hb, bounds_dev, 4 * 2 * D, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e == hipSuccess) {
        for (int i = 0; i < 2 * CUKD_MAX_DIMS; i++) tree->ibounds[i] = i < 2 * D ? (int32_t)hb[i] : 0;
        tree->nodes = out_nodes;
        tree->prim_ids = out_prims;
        tree->data = d_points;
        tree->num_prims = n;
        tree->num_nodes = num_nodes;
      }
    }
  }
  mem.release(tmp, s);
  if (e != hipSuccess) {
    if (out_nodes) mem.release(out_nodes, s);
    if (out_prims) mem.release(out_prims, s);
    set_last_error("spatial build failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  if (rc != CUKD_OK) {
    if (out_nodes) mem.release(out_nodes, s);
    if (out_prims) mem.release(out_prims, s);
  }
  return rc;
}

// ---------------------------------------------------------------------------
// spatial queries
// ---------------------------------------------------------------------------

// The stack-based query with a register-resident result keeps the first
// SP_STACK_COLS entries of its traversal stack in LDS columns: the stack
// is indexed at run time, so a per-lane array lives in scratch.  The
// 0.05-radius queries of the spatial config rarely go deeper.  (kNN-16 of
// 2M queries over 4M float4 points: 15.9 -> 12.3 ms.)  A list already in
// LDS keeps the whole stack in scratch: with both, the 32 KiB per
// workgroup cost more occupancy than the stack gained (heap kNN-16
// 14.5 -> 16.3 ms).
template <bool CCT, class R> struct SpStackCols {
  enum { value = (CCT || ResultOps<R>::kLds) ? 0 : 16 };
};

template <bool CCT, class R>
__host__ __device__ constexpr size_t sp_query_lds_words(int k, int nt)
{
  return (ResultOps<R>::kLds ? (size_t)k * 2 * nt : 0) + (size_t)SpStackCols<CCT, R>::value * 2 * nt;
}

template <int D, bool CCT, class R>
__global__ __launch_bounds__(128) void k_spatial_query(const cukd_spatial_node *nodes,
                                                       const uint32_t *prim_ids, TreeView data,
                                                       SpatialBounds bounds, const float *queries,
                                                       int nq, float cutoff, int k,
                                                       int32_t *ids, float *dist2, float *ret,
                                                       unsigned long long *stats)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  float q[D];
#pragma unroll
  for (int d = 0; d < D; d++) q[d] = queries[(size_t)i * D + d];
  HIP_DYNAMIC_SHARED(uint64_t, s_lists);
  typename ResultOps<R>::State st;
  const ListArgs no_lists{nullptr, 0, 0};
  R &res = ResultOps<R>::init(st, cutoff, k, i, no_lists,
                              LdsCol{s_lists + threadIdx.x, (int)blockDim.x});
  float lo[D], hi[D];
#pragma unroll
  for (int d = 0; d < D; d++) { lo[d] = bounds.v[d]; hi[d] = bounds.v[D + d]; }
  unsigned visits = 0;
  // stack columns: COLS node words, then as many dist2 words, each row
  // blockDim.x wide
  constexpr int COLS = SpStackCols<CCT, R>::value;
  uint32_t *stk = (uint32_t *)s_lists + threadIdx.x;
  const cukd::detail::SpatialStackCols cols{
      (int *)stk, (float *)(stk + (size_t)COLS * blockDim.x), (int)blockDim.x};
  const bool ok = cukd::detail::spatial_traverse<D, CCT, COLS>(res, q, nodes, prim_ids, lo, hi,
                                                                data, visits, cols);
  if (stats) atomicAdd(stats, (unsigned long long)visits);
  ResultOps<R>::write(st, i, k, ids, dist2, ret, no_lists);
  // stack overflow: fcp reports no result, knn returns -1 (fcp.h:317-320,
  // knn.h:620-623 print "STACK OVERFLOW" and return -1)
  if (!ok) {
    if (ret) ret[i] = -1.f;
    else ids[i] = -1;
  }
}

template <int D, bool CCT>
static hipError_t launch_spatial_m(const SpatialArgs &a, hipStream_t s)
{
  const int nt = 128, blocks = (a.nq + nt - 1) / nt;
  // dynamic LDS: k words per lane for run-time-k lists (64 KiB at k = 64),
  // or the stack-based traversal's first 16 stack entries (16 KiB)
#define CUKD_SPQ(R)                                                                          \
  hipLaunchKernelGGL((k_spatial_query<D, CCT, R>), dim3(blocks), dim3(nt),                  \
                     (sp_query_lds_words<CCT, R>(a.k, nt) * sizeof(uint32_t)), s, a.nodes,  \
                     a.prim_ids, a.data, a.bounds, a.queries, a.nq, a.cutoff, a.k, a.ids,   \
                     a.dist2, a.ret, a.stats)
  if (a.list_kind < 0) CUKD_SPQ(cukd::FCPResult);
  else if (a.list_kind == CUKD_LIST_HEAP) CUKD_SPQ(LdsHeapList);
  else if (a.k == 4) CUKD_SPQ(cukd::FixedCandidateList<4>);
  else if (a.k == 8) CUKD_SPQ(cukd::FixedCandidateList<8>);
  else if (a.k == 16) CUKD_SPQ(cukd::FixedCandidateList<16>);
  else CUKD_SPQ(DynFixedList);
#undef CUKD_SPQ
  return hipGetLastError();
}

template <int D>
static hipError_t launch_spatial_d(int method, const SpatialArgs &a, hipStream_t s)
{
  return method == CUKD_CCT ? launch_spatial_m<D, true>(a, s) : launch_spatial_m<D, false>(a, s);
}

hipError_t launch_spatial_query(int D, int method, const SpatialArgs &a, hipStream_t s)
{
  switch (D) {
    case 1: return launch_spatial_d<1>(method, a, s);
    case 2: return launch_spatial_d<2>(method, a, s);
    case 3: return launch_spatial_d<3>(method, a, s);
    case 4: return launch_spatial_d<4>(method, a, s);
    case 5: return launch_spatial_d<5>(method, a, s);
    case 6: return launch_spatial_d<6>(method, a, s);
    case 7: return launch_spatial_d<7>(method, a, s);
    default: return launch_spatial_d<8>(method, a, s);
  }
}

}  // namespace cukd_amd
