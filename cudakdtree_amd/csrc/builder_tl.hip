// builder_tl.hip -- the bitonic (N-int budget) and in-place (no temporary
// memory) builders: buildTree_bitonic / buildTree_inPlace
// (builder_bitonic.h:346-379, builder_inplace.h:444-480 in the reference).
//
// Both sort every subtree directly inside the final array.  Subtree g's
// elements live at its breadth-first positions: rank r at g*2^j + r, j the
// depth of r in the subtree.  A level's sort writes the subtree in "v
// order": sorted index v < left goes to the left child's rank v, v == left
// is the pivot at position g, v > left goes to the right child's rank
// v-left-1.  Each child's elements are therefore already in its parent's
// key order when the next level starts, so sorting by
//     (split coordinate, rank at the start of the level)
// is a stable sort -- the host builder's chain, ties included.  The sort is
// an all-ascending bitonic network (mirror step, then half-cleaners; +inf
// padding never moves) run as global compare-exchange passes for large
// strides and LDS tile passes for small ones; once subtrees fit in LDS, one
// workgroup per subtree finishes all remaining levels.
//
//  * bitonic: the rank travels with each element in an N-int array, so the
//    result equals buildTree_host for every input;
//  * in-place: nothing is allocated; ties on the split coordinate are broken
//    by the element's remaining words (a fixed total order, so the tree is
//    deterministic and valid).  It equals buildTree_host whenever the split
//    coordinates decide, e.g. on the tie-free inputs the reference's own
//    cross-builder test uses; with ties the reference's in-place builder
//    does not match its host builder either.
#include "cukd_common.h"
#include "block_sort.h"

namespace cukd_amd {

int finish_capacity_inplace(const cukd_layout &L);

// breadth-first rank <-> position inside the subtree of node g
__device__ inline int rank_pos(int g, int r) { return (g << level_of(r)) + r; }

__device__ inline int v_pos(int g, int v, int left)
{
  if (v < left) return rank_pos(2 * g + 1, v);
  if (v == left) return g;
  return rank_pos(2 * g + 2, v - left - 1);
}

__device__ inline int read_dim_field(const uint8_t *e, const cukd_layout &L)
{
  const uint8_t *p = e + L.dim_offset;
  if (L.dim_bytes == 1) return *p;
  if (L.dim_bytes == 2) return *(const uint16_t *)p;
  return *(const int32_t *)p;
}

__device__ inline void write_dim_field(uint8_t *e, const cukd_layout &L, int dim)
{
  uint8_t *p = e + L.dim_offset;
  if (L.dim_bytes == 1) *p = (uint8_t)dim;
  else if (L.dim_bytes == 2) *(uint16_t *)p = (uint16_t)dim;
  else *(int32_t *)p = dim;
}

// widestDimension on raw scalar bits (box.h:110-125): ties pick the later dim
__device__ inline int tl_widest(const uint32_t *box, int D, bool isf)
{
  int best = 0;
  if (isf) {
    float wb = 0.f;
    for (int d = 0; d < D; d++) {
      const float w = __uint_as_float(box[D + d]) - __uint_as_float(box[d]);
      if (w < wb) continue;
      wb = w;
      best = d;
    }
  } else {
    long long wb = 0;
    for (int d = 0; d < D; d++) {
      const long long w = (long long)(int)box[D + d] - (long long)(int)box[d];
      if (w < wb) continue;
      wb = w;
      best = d;
    }
  }
  return best;
}

// findBounds (builder_common.h:181-219): the domain of subtree 'node', by
// walking up through the settled ancestors
__device__ inline void find_bounds(int node, const uint32_t *world, const uint8_t *data,
                                   const cukd_layout &L, uint32_t *box)
{
  const int D = L.num_dims;
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  for (int i = 0; i < 2 * D; i++) box[i] = world[i];
  int curr = node;
  while (curr > 0) {
    const int parent = (curr + 1) / 2 - 1;
    const uint8_t *pe = data + (size_t)parent * L.stride;
    const int pd = read_dim_field(pe, L);
    const uint32_t sv = load_coord_bits(pe, L, pd);
    if (curr & 1) {
      const uint32_t u = box[D + pd];
      const bool lt = isf ? (__uint_as_float(u) < __uint_as_float(sv)) : ((int)u < (int)sv);
      box[D + pd] = lt ? u : sv;
    } else {
      const uint32_t l = box[pd];
      const bool lt = isf ? (__uint_as_float(sv) < __uint_as_float(l)) : ((int)sv < (int)l);
      box[pd] = lt ? l : sv;
    }
    curr = parent;
  }
}

// key comparison of two elements of one subtree: a > b ?
template <bool CARRY>
__device__ inline bool tl_greater(const uint8_t *ea, uint32_t ua, const uint8_t *eb, uint32_t ub,
                                  int dim, const cukd_layout &L, bool isf)
{
  const uint32_t ka = ordered_key(load_coord_bits(ea, L, dim), isf);
  const uint32_t kb = ordered_key(load_coord_bits(eb, L, dim), isf);
  if (ka != kb) return ka > kb;
  if (CARRY) return ua > ub;
  const int W = L.stride >> 2;
  for (int w = 0; w < W; w++) {
    const uint32_t a = ((const uint32_t *)ea)[w], b = ((const uint32_t *)eb)[w];
    if (a != b) return a > b;
  }
  return false;
}

__device__ inline void swap_elems(uint8_t *a, uint8_t *b, int W)
{
  uint32_t *x = (uint32_t *)a, *y = (uint32_t *)b;
  for (int w = 0; w < W; w++) {
    const uint32_t t = x[w];
    x[w] = y[w];
    y[w] = t;
  }
}

// ---------------------------------------------------------------------------
// per-level preparation: ranks (bitonic) and subtree dims (explicit)
// ---------------------------------------------------------------------------
template <bool CARRY, bool EXPLICIT>
__global__ void k_tl_prepare(int level, int n, uint8_t *data, uint32_t *rank,
                             const uint32_t *world, cukd_layout L)
{
  const int first = (1 << level) - 1;
  const int p = first + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int j = level_of(p) - level;
  const int g = ((p + 1) >> j) - 1;
  if (CARRY) rank[p] = (uint32_t)(p - (g << j));
  if (EXPLICIT) {
    uint32_t box[2 * CUKD_MAX_DIMS];
    find_bounds(g, world, data, L, box);
    write_dim_field(data + (size_t)p * L.stride, L, tl_widest(box, L.num_dims, L.scalar_type == CUKD_SCALAR_F32));
  }
}

// ---------------------------------------------------------------------------
// global compare-exchange pass: (mirror step of block size ks) or
// (half-cleaner of stride j), over all subtrees of a level
// ---------------------------------------------------------------------------
template <bool CARRY, bool EXPLICIT>
__global__ void k_tl_cx(int level, int n, int half_p, int ks, int j, int mirror, uint8_t *data,
                        uint32_t *rank, cukd_layout L)
{
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int gl = (int)(t / half_p);
  const int i = (int)(t - (long long)gl * half_p);
  if (gl >= (1 << level)) return;
  const int g = (1 << level) - 1 + gl;
  const int size = subtree_size(g, n);
  int lo, hi;
  if (mirror) {
    const int b = i / (ks >> 1), o = i - b * (ks >> 1);
    lo = b * ks + o;
    hi = b * ks + ks - 1 - o;
  } else {
    lo = 2 * i - (i & (j - 1));
    hi = lo + j;
  }
  if (hi >= size) return;
  const int left = subtree_size(2 * g + 1, n);
  const int plo = v_pos(g, lo, left), phi = v_pos(g, hi, left);
  uint8_t *ea = data + (size_t)plo * L.stride;
  uint8_t *eb = data + (size_t)phi * L.stride;
  const int dim = EXPLICIT ? read_dim_field(ea, L) : level % L.num_dims;
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const uint32_t ua = CARRY ? rank[plo] : 0u, ub = CARRY ? rank[phi] : 0u;
  if (tl_greater<CARRY>(ea, ua, eb, ub, dim, L, isf)) {
    swap_elems(ea, eb, L.stride >> 2);
    if (CARRY) { rank[plo] = ub; rank[phi] = ua; }
  }
}

// ---------------------------------------------------------------------------
// LDS tile pass: T consecutive sorted indices of one subtree; either a full
// ascending sort of the tile (first) or the half-cleaners j < T of merge
// step ks
// ---------------------------------------------------------------------------
template <bool CARRY, bool EXPLICIT, int T, int NT>
__global__ __launch_bounds__(NT) void k_tl_tile(int level, int n, int tiles_per_seg, int ks,
                                                int first, uint8_t *data, uint32_t *rank,
                                                cukd_layout L)
{
  HIP_DYNAMIC_SHARED(uint32_t, s_dyn)
  __shared__ uint16_t s_ix[T];
  __shared__ uint32_t s_rank[CARRY ? T : 1];
  const int gl = blockIdx.x / tiles_per_seg;
  const int tile = blockIdx.x - gl * tiles_per_seg;
  const int g = (1 << level) - 1 + gl;
  const int size = subtree_size(g, n);
  const int v0 = tile * T;
  if (v0 >= size) return;
  const int cnt = min(T, size - v0);
  const int left = subtree_size(2 * g + 1, n);
  const int W = L.stride >> 2;
  uint32_t *s_data = s_dyn;  // T * W words
  for (int x = threadIdx.x; x < cnt * W; x += NT) {
    const int e = x / W, w = x - e * W;
    s_data[x] = ((const uint32_t *)(data + (size_t)v_pos(g, v0 + e, left) * L.stride))[w];
  }
  for (int e = threadIdx.x; e < T; e += NT) {
    s_ix[e] = (uint16_t)e;
    if (CARRY && e < cnt) s_rank[e] = rank[v_pos(g, v0 + e, left)];
  }
  __syncthreads();
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const int dim = EXPLICIT ? read_dim_field((const uint8_t *)s_data, L) : level % L.num_dims;
  auto step = [&](int kk, int jj, bool mir) {
    for (int i = threadIdx.x; i < T / 2; i += NT) {
      int lo, hi;
      if (mir) {
        const int b = i / (kk >> 1), o = i - b * (kk >> 1);
        lo = b * kk + o;
        hi = b * kk + kk - 1 - o;
      } else {
        lo = 2 * i - (i & (jj - 1));
        hi = lo + jj;
      }
      if (hi >= cnt) continue;
      const int a = s_ix[lo], b = s_ix[hi];
      const uint8_t *ea = (const uint8_t *)(s_data + a * W);
      const uint8_t *eb = (const uint8_t *)(s_data + b * W);
      if (tl_greater<CARRY>(ea, CARRY ? s_rank[a] : 0u, eb, CARRY ? s_rank[b] : 0u, dim, L, isf)) {
        s_ix[lo] = (uint16_t)b;
        s_ix[hi] = (uint16_t)a;
      }
    }
    __syncthreads();
  };
  if (first) {
    for (int kk = 2; kk <= T; kk <<= 1) {
      step(kk, 0, true);
      for (int jj = kk >> 2; jj > 0; jj >>= 1) step(kk, jj, false);
    }
  } else {
    for (int jj = T >> 1; jj > 0; jj >>= 1) step(ks, jj, false);
  }
  // write back in the sorted order of this tile
  for (int x = threadIdx.x; x < cnt * W; x += NT) {
    const int e = x / W, w = x - e * W;
    ((uint32_t *)(data + (size_t)v_pos(g, v0 + e, left) * L.stride))[w] = s_data[s_ix[e] * W + w];
  }
  if (CARRY)
    for (int e = threadIdx.x; e < cnt; e += NT) rank[v_pos(g, v0 + e, left)] = s_rank[s_ix[e]];
}

// ---------------------------------------------------------------------------
// in-place LDS finish: one workgroup per subtree, elements staged in LDS
// ---------------------------------------------------------------------------
template <int CAP, int NT, bool EXPLICIT>
__global__ __launch_bounds__(NT) void k_tl_finish(int n, int L0, uint8_t *data,
                                                  const uint32_t *world, cukd_layout L)
{
  enum { E = CAP / NT };
  HIP_DYNAMIC_SHARED(uint32_t, s_dyn)  // CAP * W words of staged elements
  __shared__ uint32_t s_xv[CAP];
  __shared__ uint32_t s_split[EXPLICIT ? CAP : 1];
  __shared__ uint8_t s_sdim[EXPLICIT ? CAP : 1];
  __shared__ uint8_t s_bdim[EXPLICIT ? CAP / 2 : 1];
  __shared__ uint32_t s_box[2 * CUKD_MAX_DIMS];
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const int D = L.num_dims;
  const int W = L.stride >> 2;
  const int h = num_levels_for(n);
  const int hs = h - L0;
  const int Cs = 1 << hs;
  const int g_node = (1 << L0) - 1 + blockIdx.x;
  const int m = subtree_size(g_node, n);
  uint32_t *s_data = s_dyn;
  for (int x = threadIdx.x; x < m * W; x += NT) {
    const int e = x / W, w = x - e * W;
    s_data[x] = ((const uint32_t *)(data + (size_t)rank_pos(g_node, e) * L.stride))[w];
  }
  if (EXPLICIT && threadIdx.x == 0) find_bounds(g_node, world, data, L, s_box);
  const int p0 = threadIdx.x * E;
  uint64_t key[E];
  uint32_t slot[E];
#pragma unroll
  for (int e = 0; e < E; e++) slot[e] = (uint32_t)(p0 + e);
  __syncthreads();
  // every read of 'data' for this subtree is done: from here on the
  // workgroup only writes its own node slots

  for (int j = 0; j < hs; j++) {
    const int B = Cs >> j;
    const int nblk = 1 << j;
    const int first_node = (g_node + 1) * nblk - 1;
    if (EXPLICIT) {
      for (int b = threadIdx.x; b < nblk; b += NT) {
        if (first_node + b >= n) { s_bdim[b] = 0; continue; }
        uint32_t box[2 * CUKD_MAX_DIMS];
        for (int i = 0; i < 2 * D; i++) box[i] = s_box[i];
        int local = nblk - 1 + b;
        while (local > 0) {
          const int parent = (local - 1) >> 1;
          const int pd = s_sdim[parent];
          const uint32_t sv = s_split[parent];
          if (local & 1) {
            const uint32_t u = box[D + pd];
            const bool lt = isf ? (__uint_as_float(u) < __uint_as_float(sv)) : ((int)u < (int)sv);
            box[D + pd] = lt ? u : sv;
          } else {
            const uint32_t l = box[pd];
            const bool lt = isf ? (__uint_as_float(sv) < __uint_as_float(l)) : ((int)sv < (int)l);
            box[pd] = lt ? l : sv;
          }
          local = parent;
        }
        s_bdim[b] = (uint8_t)tl_widest(box, D, isf);
      }
      __syncthreads();
    }
    const int rr_dim = (L0 + j) % D;
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int p = p0 + e;
      uint64_t kk = ~0ull;
      if (p < Cs) {
        const int b = p / B;
        const int sz = subtree_size(first_node + b, n);
        if (p - b * B < sz) {
          const int d = EXPLICIT ? (int)s_bdim[b] : rr_dim;
          kk = ((uint64_t)ordered_key(load_coord_bits((const uint8_t *)(s_data + slot[e] * W), L, d), isf) << 32) | (uint64_t)p;
        }
      }
      key[e] = kk;
    }
    if (B > 1) BlockSort<CAP, NT>::sort(key, slot, B, s_xv);
#pragma unroll
    for (int e = 0; e < E; e++) {
      if (key[e] == ~0ull) continue;
      const int q = p0 + e;
      const int b = q / B;
      const int node = first_node + b;
      const int r = q - b * B;
      const int left = subtree_size(2 * node + 1, n);
      const uint32_t sl = slot[e];
      if (r == left) {
        const int d = EXPLICIT ? (int)s_bdim[b] : rr_dim;
        copy_elem(data + (size_t)node * L.stride, (const uint8_t *)(s_data + sl * W), L, d);
        if (EXPLICIT) {
          s_split[nblk - 1 + b] = load_coord_bits((const uint8_t *)(s_data + sl * W), L, d);
          s_sdim[nblk - 1 + b] = (uint8_t)d;
        }
      } else {
        const int np = r < left ? b * B + r : b * B + (B >> 1) + (r - left - 1);
        s_xv[np] = sl;
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; e++) slot[e] = s_xv[p0 + e];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
int build_bounds_only(void *d_bounds, const void *d_points, int n, const cukd_layout &L,
                      hipStream_t s);

int finish_capacity_inplace(const cukd_layout &L)
{
  // LDS bytes per element: the staged element, the sort/scatter word, and
  // for explicit dims the split value, split dim and block dim
  const int per = L.stride + 4 + (L.dim_offset >= 0 ? 6 : 0);
  int cap = 4096;
  while (cap > 512 && cap * per > 150 * 1024) cap >>= 1;
  return cap;
}

template <int CAP, int NT>
static hipError_t launch_tl_finish(bool expl, int blocks, size_t dyn, hipStream_t s, int n, int L0,
                                   uint8_t *data, const uint32_t *world, const cukd_layout &L)
{
  const void *fn = expl ? (const void *)k_tl_finish<CAP, NT, true>
                        : (const void *)k_tl_finish<CAP, NT, false>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  if (e != hipSuccess) return e;
  if (expl)
    hipLaunchKernelGGL((k_tl_finish<CAP, NT, true>), dim3(blocks), dim3(NT), dyn, s, n, L0, data,
                       world, L);
  else
    hipLaunchKernelGGL((k_tl_finish<CAP, NT, false>), dim3(blocks), dim3(NT), dyn, s, n, L0, data,
                       world, L);
  return hipGetLastError();
}

// top levels: per level, sort every subtree's T-element tiles in LDS, then
// merge them with global compare-exchange passes (strides >= T) and LDS tile
// passes (strides < T)
template <bool CARRY, bool EXPLICIT, int T>
static hipError_t tl_levels(int n, int L0, uint8_t *data, uint32_t *rank, const uint32_t *world,
                            const cukd_layout &L, hipStream_t s)
{
  enum { NT = 256 };
  const size_t dyn = (size_t)T * L.stride;
  hipError_t e = hipFuncSetAttribute((const void *)k_tl_tile<CARRY, EXPLICIT, T, NT>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn);
  if (e != hipSuccess) return e;
  for (int level = 0; level < L0; level++) {
    const int first = (1 << level) - 1;
    const int cnt = n - first;
    hipLaunchKernelGGL((k_tl_prepare<CARRY, EXPLICIT>), dim3((cnt + 255) / 256), dim3(256), 0, s,
                       level, n, data, rank, world, L);
    // the leftmost subtree of a level is the largest: it sets the padded
    // network size P
    const int maxsz = subtree_size(first, n);
    int P = 1;
    while (P < maxsz) P <<= 1;
    const int tiles = P > T ? P / T : 1;
    const int segs = 1 << level;
    hipLaunchKernelGGL((k_tl_tile<CARRY, EXPLICIT, T, NT>), dim3(segs * tiles), dim3(NT), dyn, s,
                       level, n, tiles, 0, 1, data, rank, L);
    const long long pairs = (long long)segs * (P / 2);
    const unsigned blocks = (unsigned)((pairs + 255) / 256);
    for (int ks = 2 * T; ks <= P; ks <<= 1) {
      hipLaunchKernelGGL((k_tl_cx<CARRY, EXPLICIT>), dim3(blocks), dim3(256), 0, s, level, n,
                         P / 2, ks, 0, 1, data, rank, L);
      for (int j = ks >> 2; j >= T; j >>= 1)
        hipLaunchKernelGGL((k_tl_cx<CARRY, EXPLICIT>), dim3(blocks), dim3(256), 0, s, level, n,
                           P / 2, ks, j, 0, data, rank, L);
      hipLaunchKernelGGL((k_tl_tile<CARRY, EXPLICIT, T, NT>), dim3(segs * tiles), dim3(NT), dyn,
                         s, level, n, tiles, ks, 0, data, rank, L);
    }
  }
  return hipGetLastError();
}

template <bool CARRY, bool EXPLICIT>
static hipError_t tl_levels_for_stride(int n, int L0, uint8_t *data, uint32_t *rank,
                                       const uint32_t *world, const cukd_layout &L, hipStream_t s)
{
  // tiles of at most 48 KiB of elements
  if (L.stride <= 48) return tl_levels<CARRY, EXPLICIT, 1024>(n, L0, data, rank, world, L, s);
  if (L.stride <= 96) return tl_levels<CARRY, EXPLICIT, 512>(n, L0, data, rank, world, L, s);
  return tl_levels<CARRY, EXPLICIT, 256>(n, L0, data, rank, world, L, s);
}

size_t build_tl_temp_bytes(int n, const cukd_layout &L, bool carry_rank)
{
  const int cap = finish_capacity_inplace(L);
  return carry_rank && n > cap - 1 ? (size_t)n * 4 : 0;
}

int build_tree_tl(void *d_points, int n, const cukd_layout &L, void *d_bounds, hipStream_t s,
                  const MemResource &mem, bool carry_rank)
{
  const bool expl = L.dim_offset >= 0;
  const int W = L.stride >> 2;
  uint8_t *data = (uint8_t *)d_points;
  if (d_bounds) {
    if (int rc = build_bounds_only(d_bounds, d_points, n, L, s)) return rc;
  }
  const uint32_t *world = (const uint32_t *)d_bounds;
  const int cap = finish_capacity_inplace(L);
  const int h = num_levels_for(n);
  int lg = 0;
  while ((1 << lg) < cap) lg++;
  const int L0 = h - lg > 0 ? h - lg : 0;
  uint32_t *rank = nullptr;
  if (carry_rank && L0 > 0) {
    if (mem.alloc((void **)&rank, (size_t)n * 4, s) != 0) return CUKD_ERR_ALLOC;
  }
  hipError_t e = hipSuccess;
  if (L0 > 0) {
    if (carry_rank)
      e = expl ? tl_levels_for_stride<true, true>(n, L0, data, rank, world, L, s)
               : tl_levels_for_stride<true, false>(n, L0, data, rank, world, L, s);
    else
      e = expl ? tl_levels_for_stride<false, true>(n, L0, data, rank, world, L, s)
               : tl_levels_for_stride<false, false>(n, L0, data, rank, world, L, s);
  }
  const size_t dyn = (size_t)cap * W * 4;
  const int blocks = 1 << L0;
  if (e == hipSuccess) switch (cap) {
    case 512: e = launch_tl_finish<512, 256>(expl, blocks, dyn, s, n, L0, data, world, L); break;
    case 1024: e = launch_tl_finish<1024, 256>(expl, blocks, dyn, s, n, L0, data, world, L); break;
    case 2048: e = launch_tl_finish<2048, 512>(expl, blocks, dyn, s, n, L0, data, world, L); break;
    default: e = launch_tl_finish<4096, 512>(expl, blocks, dyn, s, n, L0, data, world, L); break;
  }
  if (e == hipSuccess) e = hipGetLastError();
  if (rank) mem.release(rank, s);
  if (e != hipSuccess) {
    set_last_error("builder kernel launch failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

}  // namespace cukd_amd
