// cukd_common.h -- shared device/host helpers for the gfx950 k-d tree library.
//
// Element layout descriptor, order-preserving scalar keys, and the closed
// forms of the left-balanced tree (node levels, subtree sizes, and the
// per-level compact segment layout the builders use).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cukd_amd.h"

namespace cukd_amd {

#define CUKD_MAX_DIMS 8

// ---------------------------------------------------------------------------
// scalar keys
// ---------------------------------------------------------------------------

// Maps a coordinate's raw bits to a uint32 whose unsigned order equals the
// scalar '<' the reference comparator uses (builder_thrust.h:401-404).
// -0.0f is folded onto +0.0f because '<' treats them as equal.
__host__ __device__ inline uint32_t ordered_key(uint32_t bits, bool is_float)
{
  if (is_float) {
    if (bits == 0x80000000u) bits = 0u;
    return (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
  }
  return bits ^ 0x80000000u;
}

__host__ __device__ inline uint32_t unordered_key(uint32_t key, bool is_float)
{
  if (is_float) return (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
  return key ^ 0x80000000u;
}

// ---------------------------------------------------------------------------
// element access
// ---------------------------------------------------------------------------

__device__ inline uint32_t load_coord_bits(const uint8_t *elem,
                                           const cukd_layout &L, int d)
{
  return *(const uint32_t *)(elem + L.point_offset + 4 * d);
}

// Writes the split-dim field of an element that is being assembled word
// by word; returns the patched word for word index w.
__device__ inline uint32_t patch_dim_word(uint32_t word, int w,
                                          const cukd_layout &L, int dim)
{
  if (L.dim_offset < 0) return word;
  const int dw = L.dim_offset >> 2;
  if (w != dw) return word;
  const int shift = (L.dim_offset & 3) * 8;
  const uint32_t mask = L.dim_bytes == 4 ? 0xffffffffu
                      : (L.dim_bytes == 2 ? 0xffffu : 0xffu);
  return (word & ~(mask << shift)) | (((uint32_t)dim & mask) << shift);
}

__device__ inline void copy_elem(uint8_t *dst, const uint8_t *src,
                                 const cukd_layout &L, int dim)
{
  const int W = L.stride >> 2;
  const uint32_t *s = (const uint32_t *)src;
  uint32_t *d = (uint32_t *)dst;
  for (int w = 0; w < W; w++) d[w] = patch_dim_word(s[w], w, L, dim);
}

// ---------------------------------------------------------------------------
// left-balanced tree closed forms (helpers.h:76-178)
// ---------------------------------------------------------------------------

__host__ __device__ inline int level_of(int node)
{
#ifdef __HIP_DEVICE_COMPILE__
  return 31 - __clz(node + 1);
#else
  return 31 - __builtin_clz((unsigned)(node + 1));
#endif
}

__host__ __device__ inline int num_levels_for(int n) { return level_of(n - 1) + 1; }

// Number of nodes in the subtree under 'node' of a left-balanced tree with
// n nodes (ArbitraryBinaryTree::numNodesInSubtree, helpers.h:162-178).
__host__ __device__ inline int subtree_size(int node, int n)
{
  if (node >= n) return 0;
  const int levels_sub = num_levels_for(n) - level_of(node);
  const long long on_last = 1ll << (levels_sub - 1);
  const long long full = (1ll << levels_sub) - 1;
  long long missing = (long long)(node + 2) * on_last - 1 - n;
  missing = missing < 0 ? 0 : (missing > on_last ? on_last : missing);
  return (int)(full - missing);
}

// Compact per-level segment layout used by the builders: at level L the
// 2^L unsettled subtrees are stored back to back in node order.  Subtree
// s owns F-1 nodes on the full levels plus up to F of the 'last' deepest
// level nodes, so its offset and size are closed forms in s.
struct LevelGeom {
  int L;          // level
  int F;          // 2^(h-1-L): deepest-level slots per subtree
  int last;       // number of nodes on the deepest level
  int s_full;     // subtrees that own all F deepest-level slots
  int rem;        // deepest-level nodes of subtree s_full
  int n_seg;      // 2^L
  int n_elems;    // elements in the compact layout = n - (2^L - 1)

  __host__ __device__ inline void init(int n, int level)
  {
    const int h = num_levels_for(n);
    L = level;
    F = 1 << (h - 1 - level);
    last = n - ((1 << (h - 1)) - 1);
    s_full = last / F;
    rem = last - s_full * F;
    n_seg = 1 << level;
    n_elems = n - ((1 << level) - 1);
  }
  __host__ __device__ inline int size(int s) const
  {
    int own = last - s * F;
    own = own < 0 ? 0 : (own > F ? F : own);
    return F - 1 + own;
  }
  __host__ __device__ inline int offset(int s) const
  {
    const long long a = (long long)s * F;
    return (int)((long long)s * (F - 1) + (a < last ? a : last));
  }
  // segment containing compact position p (F >= 2)
  __host__ __device__ inline int seg_of(int p) const
  {
    const int full_sz = 2 * F - 1;
    const int full_end = s_full * full_sz;
    if (p < full_end) return p / full_sz;
    const int q = p - full_end;
    const int mid_sz = F - 1 + rem;
    if (q < mid_sz) return s_full;
    return s_full + 1 + (q - mid_sz) / (F - 1);
  }
};

// ---------------------------------------------------------------------------
// error handling
// ---------------------------------------------------------------------------

void set_last_error(const char *fmt, ...);

#define CUKD_HIP(call)                                                       \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::cukd_amd::set_last_error("%s failed: %s (%s:%d)", #call,            \
                                 hipGetErrorString(_e), __FILE__, __LINE__); \
      return CUKD_ERR_HIP;                                                   \
    }                                                                        \
  } while (0)

// Exclusive prefix sum of n uint32 (stream-ordered, several launches);
// 'scratch' holds at least scan_scratch_words(n) words.
size_t scan_scratch_words(long long n);
void scan_exclusive(const uint32_t *in, uint32_t *out, long long n, uint32_t *scratch,
                    hipStream_t s);

// Memory resource used for builder temporaries.
void clear_last_error();
int forest_num_dims(const cukd_forest *f);
int forest_num_trees(const cukd_forest *f);

struct MemResource {
  const cukd_mem_resource *user;
  int alloc(void **p, size_t bytes, hipStream_t s) const;
  void release(void *p, hipStream_t s) const;
};

}  // namespace cukd_amd
