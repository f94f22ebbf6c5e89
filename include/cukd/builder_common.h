// cukd/builder_common.h -- what the builders share (the reference's
// cukd/builder_common.h): computeBounds, and the call onto the C ABI that
// every buildTree* template makes.
#pragma once

#include "cukd/helpers.h"

namespace cukd {

namespace detail {
template <typename data_t, typename data_traits>
inline void build_via_abi(int builder, data_t *d_points, int numPoints,
                          box_t<typename data_traits::point_t> *worldBounds, hipStream_t stream,
                          GpuMemoryResource &memResource)
{
  const cukd_layout L = layout_of<data_t, data_traits>();
  cukd_mem_resource mr = bridge(memResource);
  check(cukd_build_tree(d_points, numPoints, &L, builder, worldBounds, stream, &mr), "buildTree");
}
}  // namespace detail

// World bounds of d_points[0..numPoints) into the device box d_bounds
// (builder_common.h:59-64, 151-163).
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void computeBounds(box_t<typename data_traits::point_t> *d_bounds, const data_t *d_points,
                          int numPoints, hipStream_t stream = 0)
{
  const cukd_layout L = layout_of<data_t, data_traits>();
  detail::check(cukd_compute_bounds(d_bounds, d_points, numPoints, &L, stream), "computeBounds");
}

}  // namespace cukd
