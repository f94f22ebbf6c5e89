// cukd/builder_inplace.h -- buildTree_inPlace (the reference's
// cukd/builder_inplace.h:82-102, 473-480), implemented by libcukd_amd.so.
// It allocates no memory at all (memResource is never called).  Ties on a
// split coordinate are broken by the element's bytes, so the array equals
// the host builder's whenever the split coordinates decide (always on
// tie-free input) and is otherwise a valid, deterministic tree.
#pragma once

#include "cukd/builder_common.h"

namespace cukd {

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_inPlace(data_t *d_points, int numPoints,
                              box_t<typename data_traits::point_t> *worldBounds = 0,
                              hipStream_t stream = 0,
                              GpuMemoryResource &memResource = defaultGpuMemResource())
{
  detail::build_via_abi<data_t, data_traits>(CUKD_BUILDER_ID_INPLACE, d_points, numPoints,
                                             worldBounds, stream, memResource);
}

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_inPlace(data_t *d_points, int numPoints, hipStream_t stream)
{
  buildTree_inPlace<data_t, data_traits>(d_points, numPoints, nullptr, stream);
}

// alias spelling
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_inplace(data_t *d_points, int numPoints,
                              box_t<typename data_traits::point_t> *worldBounds = 0,
                              hipStream_t stream = 0,
                              GpuMemoryResource &memResource = defaultGpuMemResource())
{
  buildTree_inPlace<data_t, data_traits>(d_points, numPoints, worldBounds, stream, memResource);
}

}  // namespace cukd
