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

enum { MORTON_BITS = 18 };

__global__ __launch_bounds__(256) void k_qbounds(uint32_t *acc, const float *q, int nq, int D)
{
  __shared__ uint32_t s_lo[CUKD_MAX_DIMS][256], s_hi[CUKD_MAX_DIMS][256];
  uint32_t lo[CUKD_MAX_DIMS], hi[CUKD_MAX_DIMS];
  for (int d = 0; d < D; d++) { lo[d] = 0xffffffffu; hi[d] = 0u; }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x)
    for (int d = 0; d < D; d++) {
      const uint32_t k = ordered_key(__float_as_uint(q[(size_t)i * D + d]), true);
      lo[d] = min(lo[d], k);
      hi[d] = max(hi[d], k);
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
    atomicMin(&acc[threadIdx.x], s_lo[threadIdx.x][0]);
    atomicMax(&acc[D + threadIdx.x], s_hi[threadIdx.x][0]);
  }
}

__global__ void k_qbounds_init(uint32_t *acc, int D)
{
  const int d = threadIdx.x;
  if (d < D) { acc[d] = 0xffffffffu; acc[D + d] = 0u; }
}

__device__ inline uint32_t morton_of(const float *q, const uint32_t *acc, int D)
{
  const int b = MORTON_BITS / D > 0 ? MORTON_BITS / D : 1;
  uint32_t code = 0;
  for (int d = 0; d < D && d < MORTON_BITS; d++) {
    const float lo = __uint_as_float(unordered_key(acc[d], true));
    const float hi = __uint_as_float(unordered_key(acc[D + d], true));
    const float w = hi - lo;
    float t = w > 0.f ? (q[d] - lo) / w : 0.f;
    int c = (int)(t * (float)(1 << b));
    c = c < 0 ? 0 : (c >= (1 << b) ? (1 << b) - 1 : c);
    for (int j = 0; j < b; j++) code |= (uint32_t)((c >> j) & 1) << (j * D + d);
  }
  return code & ((1u << MORTON_BITS) - 1);
}

__global__ void k_qhist(uint32_t *hist, const float *q, int nq, int D, const uint32_t *acc)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  atomicAdd(&hist[morton_of(q + (size_t)i * D, acc, D)], 1u);
}

__global__ void k_qscatter(uint32_t *starts, int32_t *order, const float *q, int nq, int D,
                           const uint32_t *acc)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  const uint32_t pos = atomicAdd(&starts[morton_of(q + (size_t)i * D, acc, D)], 1u);
  order[pos] = i;
}

__global__ void k_qzero(uint32_t *p, int n)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

size_t query_order_temp_bytes(int nq)
{
  const size_t buckets = (size_t)1 << MORTON_BITS;
  return ((size_t)nq * 4 + 255) / 256 * 256 + 2 * buckets * 4 + scan_scratch_words((long long)buckets) * 4 +
         2 * CUKD_MAX_DIMS * 4 + 1024;
}

// Writes the launch order of nq queries into the front of 'tmp' (which
// holds query_order_temp_bytes(nq) bytes) and returns it.
int32_t *make_query_order(void *tmp, const float *q, int nq, int D, hipStream_t s)
{
  const int buckets = 1 << MORTON_BITS;
  uint8_t *t = (uint8_t *)tmp;
  int32_t *order = (int32_t *)t;
  t += ((size_t)nq * 4 + 255) / 256 * 256;
  uint32_t *hist = (uint32_t *)t;
  t += (size_t)buckets * 4;
  uint32_t *starts = (uint32_t *)t;
  t += (size_t)buckets * 4;
  uint32_t *acc = (uint32_t *)t;
  t += 2 * CUKD_MAX_DIMS * 4;
  uint32_t *scratch = (uint32_t *)t;
  hipLaunchKernelGGL(k_qbounds_init, dim3(1), dim3(64), 0, s, acc, D);
  // few workgroups: each ends with 2*D atomics on the same 2*D words
  int blocks = (nq + 255) / 256;
  if (blocks > 128) blocks = 128;
  hipLaunchKernelGGL(k_qbounds, dim3(blocks), dim3(256), 0, s, acc, q, nq, D);
  hipLaunchKernelGGL(k_qzero, dim3(256), dim3(256), 0, s, hist, buckets);
  const int g = (nq + 255) / 256;
  hipLaunchKernelGGL(k_qhist, dim3(g), dim3(256), 0, s, hist, q, nq, D, (const uint32_t *)acc);
  scan_exclusive(hist, starts, buckets, scratch, s);
  hipLaunchKernelGGL(k_qscatter, dim3(g), dim3(256), 0, s, starts, order, q, nq, D,
                     (const uint32_t *)acc);
  return order;
}

}  // namespace cukd_amd
