// cukd/builder.h -- the reference's builder entry points (builder.h,
// builder_thrust.h, builder_bitonic.h, builder_inplace.h), implemented by
// libcukd_amd.so.  Same template signatures, HIP streams.
//
//   buildTree(points, N, worldBounds=0, stream=0, memResource=default)
//   buildTree_thrust / buildTree_bitonic / buildTree_inPlace (same args)
//   buildTree_inPlace(points, N, stream)            (builder_inplace.h:473-480)
//   buildTree_host(points, N, worldBounds=0)         (builder_thrust.h:114-118)
//   computeBounds(d_bounds, d_points, N, stream)
//
// buildTree picks its builder at compile time, as the reference's does
// (builder.h:124-150): -DCUKD_BUILDER_INPLACE selects buildTree_inPlace
// (no allocation at all), -DCUKD_BUILDER_BITONIC buildTree_bitonic, and
// otherwise buildTree_thrust runs.
//
// The points array is permuted in place into the left-balanced tree; the
// result equals the reference's buildTree_host output for any input
// without NaN coordinates (inputs with NaN are rejected).  Errors throw
// std::runtime_error, as the reference's CUKD_CUDA_CHECK does.
#pragma once

#include "cukd/builder_bitonic.h"
#include "cukd/builder_inplace.h"
#include "cukd/builder_thrust.h"

namespace cukd {

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree(data_t *d_points, int numPoints,
                      box_t<typename data_traits::point_t> *worldBounds = 0,
                      hipStream_t stream = 0,
                      GpuMemoryResource &memResource = defaultGpuMemResource())
{
#if defined(CUKD_BUILDER_INPLACE)
  buildTree_inPlace<data_t, data_traits>(d_points, numPoints, worldBounds, stream, memResource);
#elif defined(CUKD_BUILDER_BITONIC)
  buildTree_bitonic<data_t, data_traits>(d_points, numPoints, worldBounds, stream, memResource);
#else
  buildTree_thrust<data_t, data_traits>(d_points, numPoints, worldBounds, stream, memResource);
#endif
}

}  // namespace cukd
