// cukd/builder_bitonic.h -- buildTree_bitonic (the reference's
// cukd/builder_bitonic.h:83-94), implemented by libcukd_amd.so: a bitonic
// network sorts each subtree inside the final array, with n ints of
// temporary memory from memResource (builder.h:34-38).  Its array equals
// the host builder's, ties included.
#pragma once

#include "cukd/builder_common.h"

namespace cukd {

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_bitonic(data_t *d_points, int numPoints,
                              box_t<typename data_traits::point_t> *worldBounds = 0,
                              hipStream_t stream = 0,
                              GpuMemoryResource &memResource = defaultGpuMemResource())
{
  detail::build_via_abi<data_t, data_traits>(CUKD_BUILDER_ID_BITONIC, d_points, numPoints,
                                             worldBounds, stream, memResource);
}

}  // namespace cukd
