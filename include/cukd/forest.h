// cukd/forest.h -- C++ wrappers, typed by data_t / data_traits like the
// rest of the header API, for the calls without a counterpart in the
// reference: many trees built and queried together (cukd_forest_*), and
// multi-round kNN into persistent FlexHeap lists (cukd_knn_accumulate /
// cukd_knn_extract, the building block of samples/mpiHugeQuery.cu).
#pragma once

#include <vector>

#include "cukd/builder.h"
#include "cukd/fcp.h"

namespace cukd {

// Independent trees stored back to back: tree t is
// d_points[offsets[t] .. offsets[t+1]).  Every tree equals what buildTree
// builds over its points; query ids index the query's own tree.
template <typename data_t, typename data_traits = default_data_traits<data_t>>
class Forest {
 public:
  using point_t = typename data_traits::point_t;

  Forest(const std::vector<int64_t> &offsets, hipStream_t stream = 0,
         GpuMemoryResource &memResource = defaultGpuMemResource())
      : mr_(detail::bridge(memResource)), numTrees_((int)offsets.size() - 1)
  {
    const cukd_layout L = layout_of<data_t, data_traits>();
    detail::check(cukd_forest_create(&f_, offsets.data(), numTrees_, &L, stream, &mr_),
                  "Forest");
  }
  Forest(const Forest &) = delete;
  Forest &operator=(const Forest &) = delete;
  ~Forest() { (void)cukd_forest_destroy(f_, 0); }

  int numTrees() const { return numTrees_; }

  // one box_t per tree in d_worldBounds (required for explicit dims)
  void build(data_t *d_points, box_t<point_t> *d_worldBounds = nullptr, hipStream_t stream = 0)
  {
    detail::check(cukd_forest_build(f_, d_points, d_worldBounds, stream), "Forest::build");
  }

  // Allocation-free Morton-ordered queries of up to maxQueries queries.
  void reserveQueries(int maxQueries, hipStream_t stream = 0)
  {
    detail::check(cukd_forest_reserve_queries(f_, maxQueries, stream), "Forest::reserveQueries");
  }

  // Segment s = queries d_segOffsets[s] .. d_segOffsets[s+1]) on tree
  // d_segTree[s].
  void fcp(int traversal, const data_t *d_points, const box_t<point_t> *d_worldBounds,
           const point_t *d_queries, int numQueries, int numSegments, const int *d_segOffsets,
           const int *d_segTree, int *d_ids, float *d_dist2 = nullptr,
           FcpSearchParams params = FcpSearchParams{}, int flags = 0, hipStream_t stream = 0)
  {
    detail::check(cukd_forest_fcp(traversal, f_, d_points, d_worldBounds, (const float *)d_queries,
                                  numQueries, numSegments, d_segOffsets, d_segTree,
                                  params.cutOffRadius, params.eps, d_ids, d_dist2, flags, stream),
                  "Forest::fcp");
  }

  void knn(int traversal, int listKind, int k, const data_t *d_points,
           const box_t<point_t> *d_worldBounds, const point_t *d_queries, int numQueries,
           int numSegments, const int *d_segOffsets, const int *d_segTree, int *d_ids,
           float *d_dist2, float *d_maxRadius2 = nullptr, float cutOffRadius = INFINITY,
           int flags = 0, hipStream_t stream = 0)
  {
    detail::check(cukd_forest_knn(traversal, listKind, k, f_, d_points, d_worldBounds,
                                  (const float *)d_queries, numQueries, numSegments, d_segOffsets,
                                  d_segTree, cutOffRadius, d_ids, d_dist2, d_maxRadius2, flags,
                                  stream),
                  "Forest::knn");
  }

 private:
  cukd_mem_resource mr_;
  cukd_forest *f_ = nullptr;
  int numTrees_;
};

// One round of a multi-round kNN (samples/mpiHugeQuery.cu:75-88): query i
// continues its FlexHeapCandidateList at d_lists[i*k .. i*k+k); init
// fills the lists with (cutOffRadius^2, -1) first (round 0); candidate ids
// are idOffset + index in this tree.
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void knnAccumulate(int traversal, int k, const data_t *d_tree, int numPoints,
                          const box_t<typename data_traits::point_t> *d_worldBounds,
                          const typename data_traits::point_t *d_queries, int numQueries,
                          uint64_t *d_lists, bool init, float cutOffRadius = INFINITY,
                          int idOffset = 0, int flags = 0, hipStream_t stream = 0)
{
  const cukd_layout L = layout_of<data_t, data_traits>();
  detail::check(cukd_knn_accumulate(traversal, k, d_tree, numPoints, &L, d_worldBounds,
                                    (const float *)d_queries, numQueries, d_lists, init ? 1 : 0,
                                    cutOffRadius, idOffset, flags, stream, nullptr),
                "knnAccumulate");
}

// extractFinalResult (samples/mpiHugeQuery.cu:90-103): d_radius[i] =
// sqrt of list i's max dist2; d_ids / d_dist2 (nullable) get the entries
// in ascending order.
inline void knnExtract(const uint64_t *d_lists, int numQueries, int k, int *d_ids,
                       float *d_dist2, float *d_radius, hipStream_t stream = 0)
{
  detail::check(cukd_knn_extract(d_lists, numQueries, k, d_ids, d_dist2, d_radius, stream),
                "knnExtract");
}

}  // namespace cukd
