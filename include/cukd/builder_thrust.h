// cukd/builder_thrust.h -- buildTree_thrust and buildTree_host (the
// reference's cukd/builder_thrust.h:90-118), implemented by libcukd_amd.so.
//
// There is no thrust underneath: buildTree_thrust runs the library's
// radix-select builder, whose array equals the reference host builder's
// (the stable per-level sort of builder_thrust.h:514-601) for every input
// without NaN coordinates, ties included.  buildTree_host builds the same
// tree over host (or managed) memory.
#pragma once

#include "cukd/builder_common.h"

namespace cukd {

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_thrust(data_t *d_points, int numPoints,
                             box_t<typename data_traits::point_t> *worldBounds = 0,
                             hipStream_t stream = 0,
                             GpuMemoryResource &memResource = defaultGpuMemResource())
{
  detail::build_via_abi<data_t, data_traits>(CUKD_BUILDER_ID_THRUST, d_points, numPoints,
                                             worldBounds, stream, memResource);
}

// Builds over host-accessible points; worldBounds, when non-null, is a host
// box that receives the bounds (required for explicit dims, as at
// builder_thrust.h:567-571).  Returns when the array is final.
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_host(data_t *points, int numPoints,
                           box_t<typename data_traits::point_t> *worldBounds = 0)
{
  const cukd_layout L = layout_of<data_t, data_traits>();
  detail::check(cukd_build_tree_host(points, numPoints, &L, worldBounds), "buildTree_host");
}

}  // namespace cukd
