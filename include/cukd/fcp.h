// cukd/fcp.h -- find-closest-point queries (drop-in for the reference's
// cukd/fcp.h): FcpSearchParams, FCPResult, and stackBased / stackFree /
// cct ::fcp as device-inline functions for user kernels, plus a batched
// host call (cukd::batch::fcp) that runs one query per lane through the
// library's kernels.
#pragma once

#include <limits.h>
#include <math.h>

#include "cukd/traverse.h"

namespace cukd {

struct FcpSearchParams {
  // kept for source compatibility; every traversal ignores it, as in the
  // reference
  int far_node_inspect_budget = INT_MAX;
  // only points strictly closer than this radius are reported
  float cutOffRadius = INFINITY;
  // stack-free traversal only: descend into a far child when
  // (1+eps) * planeDist^2 < current cull distance
  float eps = 0;
};

// closest candidate so far; the first strictly closer candidate wins
struct FCPResult {
  inline __both__ float initialCullDist2() const { return closestDist2; }
  inline __both__ float clear(float initialDist2)
  {
    closestDist2 = initialDist2;
    closestPrimID = -1;
    return closestDist2;
  }
  inline __both__ float processCandidate(int candPrimID, float candDist2)
  {
    if (candDist2 < closestDist2) {
      closestDist2 = candDist2;
      closestPrimID = candPrimID;
    }
    return closestDist2;
  }
  inline __both__ int returnValue() const { return closestPrimID; }

  int closestPrimID;
  float closestDist2;
};

namespace stackBased {
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ int fcp(typename data_traits::point_t queryPoint, const data_t *dataPoints,
                        int numDataPoints, FcpSearchParams params = FcpSearchParams{})
{
  FCPResult r;
  r.clear(sqr(params.cutOffRadius));
  traverse_default<FCPResult, data_t, data_traits>(r, queryPoint, dataPoints, numDataPoints);
  return r.returnValue();
}
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ int fcp(typename data_traits::point_t queryPoint,
                        const box_t<typename data_traits::point_t> /*worldBounds*/,
                        const data_t *dataPoints, int numDataPoints,
                        FcpSearchParams params = FcpSearchParams{})
{
  return fcp<data_t, data_traits>(queryPoint, dataPoints, numDataPoints, params);
}
}  // namespace stackBased

namespace stackFree {
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ int fcp(typename data_traits::point_t queryPoint, const data_t *dataPoints,
                        int numDataPoints, FcpSearchParams params = FcpSearchParams{})
{
  FCPResult r;
  r.clear(sqr(params.cutOffRadius));
  traverse_stack_free<FCPResult, data_t, data_traits>(r, queryPoint, dataPoints, numDataPoints,
                                                      params.eps);
  return r.returnValue();
}
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ int fcp(typename data_traits::point_t queryPoint,
                        const box_t<typename data_traits::point_t> /*worldBounds*/,
                        const data_t *dataPoints, int numDataPoints,
                        FcpSearchParams params = FcpSearchParams{})
{
  return fcp<data_t, data_traits>(queryPoint, dataPoints, numDataPoints, params);
}
}  // namespace stackFree

namespace cct {
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ int fcp(typename data_traits::point_t queryPoint,
                        const box_t<typename data_traits::point_t> worldBounds,
                        const data_t *dataPoints, int numDataPoints,
                        FcpSearchParams params = FcpSearchParams{})
{
  FCPResult r;
  r.clear(sqr(params.cutOffRadius));
  traverse_cct<FCPResult, data_t, data_traits>(r, queryPoint, worldBounds, dataPoints,
                                               numDataPoints);
  return r.returnValue();
}
}  // namespace cct

// ---------------------------------------------------------------------------
// batched host call: d_ids[i] (and d_dist2[i] if non-null) for every query
// ---------------------------------------------------------------------------
namespace batch {
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void fcp(int traversal, const typename data_traits::point_t *d_queries, int numQueries,
                const data_t *d_nodes, int numNodes,
                const box_t<typename data_traits::point_t> *d_worldBounds, int *d_ids,
                float *d_dist2 = nullptr, FcpSearchParams params = FcpSearchParams{},
                hipStream_t stream = 0, int flags = 0)
{
  static_assert(sizeof(typename data_traits::point_t) ==
                    4 * point_traits<typename data_traits::point_t>::num_dims,
                "queries must be packed float points");
  const cukd_layout L = layout_of<data_t, data_traits>();
  detail::check(cukd_fcp(traversal, d_nodes, numNodes, &L, d_worldBounds,
                         (const float *)d_queries, numQueries, params.cutOffRadius, params.eps,
                         d_ids, d_dist2, flags, stream, nullptr),
                "batch::fcp");
}
}  // namespace batch

}  // namespace cukd
