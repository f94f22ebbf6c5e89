// qorder.hip -- Morton ordering of query batches (CUKD_QUERY_SORT).
//
// Queries are bucketed by an 18-bit Morton code of their position inside
// the batch's bounding box (bits spread evenly over the dims), so the lanes
// of a wave start from nearby points and walk mostly the same tree nodes.
// Only the launch order changes: every query's result is written at its own
// index, so results are identical with and without ordering.  Order inside a
// bucket is whatever the atomics give; it affects speed, never results.
#include "cukd_common.h"

namespace cukd_amd {

enum { MORTON_BITS = 18, QB_SLOTS = 32, QB_NT = 256 };

// Bounds accumulators: QB_SLOTS copies of (lo[CUKD_MAX_DIMS], hi[CUKD_MAX_DIMS])
// as ordered keys; workgroup b folds into copy b % QB_SLOTS, so each word
// takes few atomics, and readers fold the copies.
enum { QB_WORDS = QB_SLOTS * 2 * CUKD_MAX_DIMS };

// clears the histogram and sets the accumulators to empty bounds
__global__ void k_qinit(uint32_t *hist, int buckets, uint32_t *acc)
{
  const int stride = gridDim.x * blockDim.x;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < buckets; i += stride) hist[i] = 0u;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < QB_WORDS; i += stride)
    acc[i] = (i % (2 * CUKD_MAX_DIMS)) < CUKD_MAX_DIMS ? 0xffffffffu : 0u;
}

// Per-dim min / max of the nq*D query coordinates.  The flat array is read
// coalesced with a thread stride T that is a multiple of D, so each thread
// sees one dim only; four loads are in flight per thread.
__global__ __launch_bounds__(QB_NT) void k_qbounds(uint32_t *acc, const float *q, long long E, int D,
                                                  long long T)
{
  __shared__ uint32_t s_lo[CUKD_MAX_DIMS], s_hi[CUKD_MAX_DIMS];
  if (threadIdx.x < CUKD_MAX_DIMS) {
    s_lo[threadIdx.x] = 0xffffffffu;
    s_hi[threadIdx.x] = 0u;
  }
  __syncthreads();
  const long long tid = (long long)blockIdx.x * QB_NT + threadIdx.x;
  if (tid < T) {
    uint32_t lo = 0xffffffffu, hi = 0u;
    for (long long e = tid; e < E; e += 4 * T) {
      uint32_t v[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const long long x = e + j * T;
        v[j] = x < E ? ordered_key(__float_as_uint(q[x]), true) : ordered_key(__float_as_uint(q[e]), true);
      }
#pragma unroll
      for (int j = 0; j < 4; j++) {
        lo = min(lo, v[j]);
        hi = max(hi, v[j]);
      }
    }
    const int d = (int)(tid % D);
    if (tid < E) {
      atomicMin(&s_lo[d], lo);
      atomicMax(&s_hi[d], hi);
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < D) {
    uint32_t *a = acc + (blockIdx.x % QB_SLOTS) * 2 * CUKD_MAX_DIMS;
    if (s_lo[threadIdx.x] != 0xffffffffu) atomicMin(&a[threadIdx.x], s_lo[threadIdx.x]);
    if (s_hi[threadIdx.x] != 0u) atomicMax(&a[CUKD_MAX_DIMS + threadIdx.x], s_hi[threadIdx.x]);
  }
}

// the batch box, folded from the accumulators (written by the previous
// kernel, so plain loads see them): thread t < D folds lower bound t over
// the QB_SLOTS copies, thread D + t upper bound t
__device__ inline void qbox_load(const uint32_t *acc, int D, float *s_lo, float *s_hi)
{
  const int t = (int)threadIdx.x;
  if (t < 2 * D) {
    const bool up = t >= D;
    const int w0 = up ? CUKD_MAX_DIMS + (t - D) : t;
    uint32_t v = up ? 0u : 0xffffffffu;
#pragma unroll 8
    for (int c = 0; c < QB_SLOTS; c++) {
      const uint32_t x = acc[c * 2 * CUKD_MAX_DIMS + w0];
      v = up ? max(v, x) : min(v, x);
    }
    const float f = __uint_as_float(unordered_key(v, true));
    if (up) s_hi[t - D] = f;
    else s_lo[t] = f;
  }
  __syncthreads();
}

__device__ inline uint32_t morton_of(const float *q, const float *lo_, const float *hi_, int D)
{
  const int b = MORTON_BITS / D > 0 ? MORTON_BITS / D : 1;
  uint32_t code = 0;
  for (int d = 0; d < D && d < MORTON_BITS; d++) {
    const float lo = lo_[d];
    const float hi = hi_[d];
    const float w = hi - lo;
    float t = w > 0.f ? (q[d] - lo) / w : 0.f;
    int c = (int)(t * (float)(1 << b));
    c = c < 0 ? 0 : (c >= (1 << b) ? (1 << b) - 1 : c);
    for (int j = 0; j < b; j++) code |= (uint32_t)((c >> j) & 1) << (j * D + d);
  }
  return code & ((1u << MORTON_BITS) - 1);
}

// bucket counts; each query keeps its bucket and its rank inside it
__global__ __launch_bounds__(256) void k_qhist(uint32_t *hist, uint2 *code_rank, const float *q,
                                              int nq, int D, const uint32_t *acc)
{
  __shared__ float s_lo[CUKD_MAX_DIMS], s_hi[CUKD_MAX_DIMS];
  qbox_load(acc, D, s_lo, s_hi);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const uint32_t c = morton_of(q + (size_t)i * D, s_lo, s_hi, D);
  code_rank[i] = make_uint2(c, atomicAdd(&hist[c], 1u));
}

__global__ void k_qscatter(const uint32_t *starts, int32_t *order, const uint2 *code_rank, int nq)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const uint2 cr = code_rank[i];
  order[starts[cr.x] + cr.y] = i;
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

size_t query_order_temp_bytes(int nq)
{
  const size_t buckets = (size_t)1 << MORTON_BITS;
  return align256((size_t)nq * 4) + align256((size_t)nq * 8) + 2 * buckets * 4 +
         align256(scan_scratch_words((long long)buckets) * 4) + QB_WORDS * 4 + 256;
}

// Writes the launch order of nq queries into the front of 'tmp' (which
// holds query_order_temp_bytes(nq) bytes) and returns it.
int32_t *make_query_order(void *tmp, const float *q, int nq, int D, hipStream_t s)
{
  const int buckets = 1 << MORTON_BITS;
  uint8_t *t = (uint8_t *)tmp;
  int32_t *order = (int32_t *)t;
  if (nq <= 0) return order;
  t += align256((size_t)nq * 4);
  uint2 *code_rank = (uint2 *)t;
  t += align256((size_t)nq * 8);
  uint32_t *hist = (uint32_t *)t;
  t += (size_t)buckets * 4;
  uint32_t *starts = (uint32_t *)t;
  t += (size_t)buckets * 4;
  uint32_t *scratch = (uint32_t *)t;
  t += align256(scan_scratch_words((long long)buckets) * 4);
  uint32_t *acc = (uint32_t *)t;
  hipLaunchKernelGGL(k_qinit, dim3(256), dim3(256), 0, s, hist, buckets, acc);
  const long long E = (long long)nq * D;
  long long blocks = (E + QB_NT * 4 - 1) / (QB_NT * 4);
  if (blocks > 1024) blocks = 1024;
  long long T = blocks * QB_NT;
  T -= T % D;
  if (T < D) T = D;  // E >= D, and blocks * QB_NT >= 256 >= D
  hipLaunchKernelGGL(k_qbounds, dim3((unsigned)blocks), dim3(QB_NT), 0, s, acc, q, E, D, T);
  const int g = (nq + 255) / 256;
  hipLaunchKernelGGL(k_qhist, dim3(g), dim3(256), 0, s, hist, code_rank, q, nq, D,
                     (const uint32_t *)acc);
  scan_exclusive(hist, starts, buckets, scratch, s);
  hipLaunchKernelGGL(k_qscatter, dim3(g), dim3(256), 0, s, (const uint32_t *)starts, order,
                     (const uint2 *)code_rank, nq);
  return order;
}

}  // namespace cukd_amd
