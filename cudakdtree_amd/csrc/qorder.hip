// qorder.hip -- Morton ordering of query batches (CUKD_QUERY_SORT).
//
// Queries are bucketed by a Morton code of their position inside the
// batch's bounding box (MORTON_BITS bits spread evenly over the dims), so
// the lanes of a wave start from nearby points and walk mostly the same
// tree nodes.  Only the launch order changes: every query's result is
// written at its own index, so results are identical with and without
// ordering.  Order inside a cell is whatever the atomics give; it affects
// speed, never results.
//
// A counting sort without returning global atomics: each tile of
// QO_TILE queries counts its cells in LDS (k_qcount), one thread per cell
// turns the per-tile counts into per-tile starts and a cell total
// (k_qcolscan), and each tile then places its queries from its starts with
// LDS atomics (k_qplace), after scanning the cell totals itself.
#include "cukd_common.h"

namespace cukd_amd {

// 15-bit cells: on a 10M-point tree, kNN-8 over 1M uniform queries took
// 0.789 / 0.764 / 0.747 / 0.737 ms launched in 9 / 12 / 15 / 18-bit cell
// order; 2^15 counters (128 KB) are the most one workgroup's LDS holds.
enum { MORTON_BITS = 15, QB_SLOTS = 32, QB_NT = 256 };
enum { QO_CELLS = 1 << MORTON_BITS, QO_NT = 1024, QO_ITEMS = 16, QO_TILE = QO_NT * QO_ITEMS };

// Bounds accumulators: QB_SLOTS copies of (lo[CUKD_MAX_DIMS], hi[CUKD_MAX_DIMS])
// as ordered keys; workgroup b folds into copy b % QB_SLOTS, so each word
// takes few atomics, and readers fold the copies.
enum { QB_WORDS = QB_SLOTS * 2 * CUKD_MAX_DIMS };

// sets the accumulators to empty bounds
__global__ void k_qinit(uint32_t *acc)
{
  for (int i = threadIdx.x; i < QB_WORDS; i += blockDim.x)
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

// Morton cell of one query from its coordinates (DC dims)
template <int DC>
__device__ inline uint32_t morton_cell(const float *x, const float *lo_, const float *hi_)
{
  constexpr int b = MORTON_BITS / DC > 0 ? MORTON_BITS / DC : 1;
  uint32_t code = 0;
#pragma unroll
  for (int d = 0; d < DC && d < MORTON_BITS; d++) {
    const float lo = lo_[d];
    const float hi = hi_[d];
    const float w = hi - lo;
    float t = w > 0.f ? (x[d] - lo) / w : 0.f;
    int c = (int)(t * (float)(1 << b));
    c = c < 0 ? 0 : (c >= (1 << b) ? (1 << b) - 1 : c);
#pragma unroll
    for (int j = 0; j < b; j++) code |= (uint32_t)((c >> j) & 1) << (j * DC + d);
  }
  return code & ((1u << MORTON_BITS) - 1);
}

// The cells of a thread's queries in one tile, QO_BATCH at a time: every
// coordinate load of a batch is issued before any is used.
enum { QO_BATCH = 4, QO_NONE = 0xffffffffu };
template <int DC>
__device__ inline void tile_cells(uint32_t (&cell)[QO_ITEMS], const float *q, int nq,
                                  const float *s_lo, const float *s_hi)
{
  const long long i0 = (long long)blockIdx.x * QO_TILE + threadIdx.x;
#pragma unroll
  for (int k0 = 0; k0 < QO_ITEMS; k0 += QO_BATCH) {
    float x[QO_BATCH][DC];
#pragma unroll
    for (int k = 0; k < QO_BATCH; k++) {
      const long long i = i0 + (long long)(k0 + k) * QO_NT;
#pragma unroll
      for (int d = 0; d < DC; d++) x[k][d] = i < nq ? q[(size_t)i * DC + d] : 0.f;
    }
#pragma unroll
    for (int k = 0; k < QO_BATCH; k++) {
      const long long i = i0 + (long long)(k0 + k) * QO_NT;
      cell[k0 + k] = i < nq ? morton_cell<DC>(x[k], s_lo, s_hi) : (uint32_t)QO_NONE;
    }
  }
}

// Cell counts of one tile of QO_TILE queries: row 'tile' of th (tile-major,
// QO_CELLS words a row).
template <int DC>
__global__ __launch_bounds__(QO_NT) void k_qcount(uint32_t *th, const float *q, int nq,
                                                 const uint32_t *acc)
{
  __shared__ uint32_t h[QO_CELLS];
  __shared__ float s_lo[CUKD_MAX_DIMS], s_hi[CUKD_MAX_DIMS];
  for (int c = threadIdx.x; c < QO_CELLS; c += QO_NT) h[c] = 0u;
  qbox_load(acc, DC, s_lo, s_hi);  // ends with a barrier
  uint32_t cell[QO_ITEMS];
  tile_cells<DC>(cell, q, nq, s_lo, s_hi);
#pragma unroll
  for (int k = 0; k < QO_ITEMS; k++)
    if (cell[k] != QO_NONE) atomicAdd(&h[cell[k]], 1u);
  __syncthreads();
  uint32_t *row = th + (size_t)blockIdx.x * QO_CELLS;
  for (int c = threadIdx.x; c < QO_CELLS; c += QO_NT) row[c] = h[c];
}

// One thread per cell: the counts of the tiles become each tile's start
// inside the cell (exclusive, in tile order), and tot[cell] the cell's
// size.  Eight tiles' counts are loaded before any is rewritten.
__global__ void k_qcolscan(uint32_t *th, uint32_t *tot, int ntiles)
{
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= QO_CELLS) return;
  uint32_t run = 0;
  for (int t0 = 0; t0 < ntiles; t0 += 8) {
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; j++)
      v[j] = t0 + j < ntiles ? th[(size_t)(t0 + j) * QO_CELLS + c] : 0u;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      if (t0 + j < ntiles) th[(size_t)(t0 + j) * QO_CELLS + c] = run;
      run += v[j];
    }
  }
  tot[c] = run;
}

// Places one tile's queries: the cell starts are the exclusive scan of the
// cell totals plus the tile's start inside each cell; queries take
// consecutive slots from there.
template <int DC>
__global__ __launch_bounds__(QO_NT) void k_qplace(int32_t *order, const uint32_t *th,
                                                 const uint32_t *tot, const float *q, int nq,
                                                 const uint32_t *acc)
{
  enum { CPT = QO_CELLS / QO_NT };  // cells per thread, consecutive
  static_assert(QO_CELLS % QO_NT == 0, "cells per thread");
  __shared__ uint32_t base[QO_CELLS];
  __shared__ uint32_t wsum[QO_NT / 64];
  __shared__ float s_lo[CUKD_MAX_DIMS], s_hi[CUKD_MAX_DIMS];
  const int tid = (int)threadIdx.x, lane = tid & 63, w = tid >> 6;
  uint32_t v[CPT], sum = 0;
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    v[j] = tot[tid * CPT + j];
    sum += v[j];
  }
  // exclusive scan of the per-thread sums over the workgroup
  uint32_t incl = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)incl, o);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint32_t run = incl - sum;
  for (int i = 0; i < w; i++) run += wsum[i];
  const uint32_t *row = th + (size_t)blockIdx.x * QO_CELLS;
#pragma unroll
  for (int j = 0; j < CPT; j++) {
    base[tid * CPT + j] = run + row[tid * CPT + j];
    run += v[j];
  }
  qbox_load(acc, DC, s_lo, s_hi);  // ends with a barrier
  uint32_t cell[QO_ITEMS];
  tile_cells<DC>(cell, q, nq, s_lo, s_hi);
  const long long i0 = (long long)blockIdx.x * QO_TILE + threadIdx.x;
#pragma unroll
  for (int k = 0; k < QO_ITEMS; k++)
    if (cell[k] != QO_NONE) order[atomicAdd(&base[cell[k]], 1u)] = (int32_t)(i0 + (long long)k * QO_NT);
}

static size_t align256(size_t b) { return (b + 255) / 256 * 256; }

static int order_tiles(int nq) { return (nq + QO_TILE - 1) / QO_TILE; }

size_t query_order_temp_bytes(int nq)
{
  return align256((size_t)nq * 4) + align256((size_t)order_tiles(nq) * QO_CELLS * 4) +
         align256((size_t)QO_CELLS * 4) + QB_WORDS * 4 + 256;
}

// Writes the launch order of nq queries into the front of 'tmp' (which
// holds query_order_temp_bytes(nq) bytes) and returns it.
int32_t *make_query_order(void *tmp, const float *q, int nq, int D, hipStream_t s)
{
  uint8_t *t = (uint8_t *)tmp;
  int32_t *order = (int32_t *)t;
  if (nq <= 0) return order;
  const int tiles = order_tiles(nq);
  t += align256((size_t)nq * 4);
  uint32_t *th = (uint32_t *)t;
  t += align256((size_t)tiles * QO_CELLS * 4);
  uint32_t *tot = (uint32_t *)t;
  t += align256((size_t)QO_CELLS * 4);
  uint32_t *acc = (uint32_t *)t;
  hipLaunchKernelGGL(k_qinit, dim3(1), dim3(256), 0, s, acc);
  const long long E = (long long)nq * D;
  long long blocks = (E + QB_NT * 4 - 1) / (QB_NT * 4);
  if (blocks > 1024) blocks = 1024;
  long long T = blocks * QB_NT;
  T -= T % D;
  if (T < D) T = D;  // E >= D, and blocks * QB_NT >= 256 >= D
  hipLaunchKernelGGL(k_qbounds, dim3((unsigned)blocks), dim3(QB_NT), 0, s, acc, q, E, D, T);
#define CUKD_QORDER_CASE(DCv)                                                                  \
  case DCv:                                                                                    \
    hipLaunchKernelGGL(k_qcount<DCv>, dim3(tiles), dim3(QO_NT), 0, s, th, q, nq,              \
                       (const uint32_t *)acc);                                                 \
    hipLaunchKernelGGL(k_qcolscan, dim3((QO_CELLS + 255) / 256), dim3(256), 0, s, th, tot,    \
                       tiles);                                                                 \
    hipLaunchKernelGGL(k_qplace<DCv>, dim3(tiles), dim3(QO_NT), 0, s, order,                  \
                       (const uint32_t *)th, (const uint32_t *)tot, q, nq,                     \
                       (const uint32_t *)acc);                                                 \
    break
  switch (D) {
    CUKD_QORDER_CASE(1);
    CUKD_QORDER_CASE(2);
    CUKD_QORDER_CASE(3);
    CUKD_QORDER_CASE(4);
    CUKD_QORDER_CASE(5);
    CUKD_QORDER_CASE(6);
    CUKD_QORDER_CASE(7);
    default: CUKD_QORDER_CASE(8);
  }
#undef CUKD_QORDER_CASE
  return order;
}

}  // namespace cukd_amd
