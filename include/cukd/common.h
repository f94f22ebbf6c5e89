// cukd/common.h -- definitions every cukd header shares (drop-in for the
// reference's cukd/common.h): the __both__ qualifier, divRoundUp, and the
// traversal statistics switch.
//
// Traversal statistics (common.h:345-360 of the reference): compile a user
// translation unit with -DCUKD_ENABLE_STATS=1 and point
// cukd::g_traversalStats at a device counter, e.g.
//
//   unsigned long long *d_stats;            // device memory, zeroed
//   hipMemcpyToSymbol(HIP_SYMBOL(cukd::g_traversalStats), &d_stats,
//                     sizeof(d_stats));
//
// Every device-side traversal in that translation unit (traverse_default,
// traverse_stack_free, traverse_cct, the fcp/knn wrappers and the spatial
// queries) then adds the nodes it visited to *d_stats, counted as the
// reference counts them (traverse-default-stack-based.h:64,
// traverse-stack-free.h:58, traverse-cct.h:71, fcp.h:294,327).  It adds
// them once per query, not once per node.  The variable is per
// translation unit, as HIP device variables are without relocatable
// device code; the library's own batch launches count through
// cukd_set_traversal_stats instead.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <stdexcept>

#ifndef __both__
#define __both__ __host__ __device__
#endif

#if defined(CUKD_ENABLE_STATS) && defined(__HIP_DEVICE_COMPILE__)
#define CUKD_STATS(a) a
#define CUKD_STATS_ARG(a, b) a,
#else
#define CUKD_STATS(a) /* nothing */
#define CUKD_STATS_ARG(a, b) /* nothing */
#endif

#if CUKD_ENABLE_STATS
namespace cukd {
static __constant__ __device__ unsigned long long *g_traversalStats;
}
#endif

// HIP error checks for callers, in the manner of the reference's
// CUKD_CUDA_CHECK / CUKD_CUDA_CALL / CUKD_CUDA_SYNC_CHECK (common.h:276-314):
// print the failing call and throw std::runtime_error.
#define CUKD_HIP_CHECK(call)                                                         \
  do {                                                                               \
    const hipError_t cukd_rc_ = (call);                                              \
    if (cukd_rc_ != hipSuccess) {                                                    \
      fprintf(stderr, "HIP call (%s) failed with code %d (line %d): %s\n", #call,    \
              (int)cukd_rc_, __LINE__, hipGetErrorString(cukd_rc_));                \
      throw std::runtime_error("fatal hip error");                                   \
    }                                                                                \
  } while (0)
#define CUKD_HIP_CALL(call) CUKD_HIP_CHECK(hip##call)
#define CUKD_HIP_SYNC_CHECK()                \
  do {                                       \
    CUKD_HIP_CHECK(hipDeviceSynchronize());  \
    CUKD_HIP_CHECK(hipGetLastError());       \
  } while (0)

namespace cukd {

inline __both__ int32_t divRoundUp(int32_t a, int32_t b) { return (a + b - 1) / b; }
inline __both__ uint32_t divRoundUp(uint32_t a, uint32_t b) { return (a + b - 1) / b; }
inline __both__ int64_t divRoundUp(int64_t a, int64_t b) { return (a + b - 1) / b; }

namespace detail {
// adds one query's visit count to the statistics counter (no-op unless
// CUKD_ENABLE_STATS)
inline __both__ void count_visits(unsigned visits)
{
  (void)visits;
  CUKD_STATS(if (cukd::g_traversalStats)::atomicAdd(cukd::g_traversalStats,
                                                      (unsigned long long)visits));
}
}  // namespace detail

}  // namespace cukd
