// multiround.hip -- final extraction of persistent multi-round kNN lists
// (extractFinalResult, samples/mpiHugeQuery.cu:90-103).
#pragma clang fp contract(off)

#include "cukd_common.h"
#include "query_impl.h"

namespace cukd_amd {

// One query per lane: radius = sqrt(max dist2) unless infinite, as the
// reference; ids/dist2 (each nullable) get the k entries in ascending
// (dist2, id) order.
__global__ __launch_bounds__(256) void k_knn_extract(const uint64_t *lists, int nq, int k,
                                                     int32_t *ids, float *dist2, float *radius)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const uint64_t *src = lists + (size_t)q * k;
  const uint64_t top = src[0];
  if (radius) {
    const float m = __uint_as_float((uint32_t)(top >> 32));
    radius[q] = isinf(m) ? m : sqrtf(m);
  }
  if (!ids && !dist2) return;
  uint64_t e[64];
  for (int i = 0; i < k; i++) {
    const uint64_t v = src[i];
    int j = i;
    while (j > 0 && e[j - 1] > v) { e[j] = e[j - 1]; j--; }
    e[j] = v;
  }
  for (int i = 0; i < k; i++) {
    if (ids) ids[(size_t)q * k + i] = (int32_t)(uint32_t)e[i];
    if (dist2) dist2[(size_t)q * k + i] = __uint_as_float((uint32_t)(e[i] >> 32));
  }
}

hipError_t launch_knn_extract(const uint64_t *lists, int nq, int k, int32_t *ids, float *dist2,
                              float *radius, hipStream_t s)
{
  hipLaunchKernelGGL(k_knn_extract, dim3((nq + 255) / 256), dim3(256), 0, s, lists, nq, k, ids,
                     dist2, radius);
  return hipGetLastError();
}

}  // namespace cukd_amd
