// cukd/builder.h -- the reference's builder entry points (builder.h,
// builder_thrust.h, builder_bitonic.h, builder_inplace.h), implemented by
// libcukd_amd.so.  Same template signatures, HIP streams.
//
//   buildTree(points, N, worldBounds=0, stream=0, memResource=default)
//   buildTree_thrust / buildTree_bitonic / buildTree_inPlace (same args)
//   buildTree_inPlace(points, N, stream)            (builder_inplace.h:473-480)
//   computeBounds(d_bounds, d_points, N, stream)
//
// The points array is permuted in place into the left-balanced tree; the
// result equals the reference's buildTree_host output for any input
// without NaN coordinates.  Errors throw std::runtime_error, as the
// reference's CUKD_CUDA_CHECK does.
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

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree(data_t *d_points, int numPoints,
                      box_t<typename data_traits::point_t> *worldBounds = 0,
                      hipStream_t stream = 0,
                      GpuMemoryResource &memResource = defaultGpuMemResource())
{
  detail::build_via_abi<data_t, data_traits>(CUKD_BUILDER_DEFAULT, d_points, numPoints,
                                             worldBounds, stream, memResource);
}

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_thrust(data_t *d_points, int numPoints,
                             box_t<typename data_traits::point_t> *worldBounds = 0,
                             hipStream_t stream = 0,
                             GpuMemoryResource &memResource = defaultGpuMemResource())
{
  detail::build_via_abi<data_t, data_traits>(CUKD_BUILDER_THRUST, d_points, numPoints,
                                             worldBounds, stream, memResource);
}

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_bitonic(data_t *d_points, int numPoints,
                              box_t<typename data_traits::point_t> *worldBounds = 0,
                              hipStream_t stream = 0,
                              GpuMemoryResource &memResource = defaultGpuMemResource())
{
  detail::build_via_abi<data_t, data_traits>(CUKD_BUILDER_BITONIC, d_points, numPoints,
                                             worldBounds, stream, memResource);
}

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree_inPlace(data_t *d_points, int numPoints,
                              box_t<typename data_traits::point_t> *worldBounds = 0,
                              hipStream_t stream = 0,
                              GpuMemoryResource &memResource = defaultGpuMemResource())
{
  detail::build_via_abi<data_t, data_traits>(CUKD_BUILDER_INPLACE, d_points, numPoints,
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

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void computeBounds(box_t<typename data_traits::point_t> *d_bounds, const data_t *d_points,
                          int numPoints, hipStream_t stream = 0)
{
  const cukd_layout L = layout_of<data_t, data_traits>();
  detail::check(cukd_compute_bounds(d_bounds, d_points, numPoints, &L, stream), "computeBounds");
}

}  // namespace cukd
