// cukd/spatial-kdtree.h -- spatial k-d tree (drop-in for the reference's
// cukd/spatial-kdtree.h and its queries in fcp.h / knn.h).
//
//   SpatialKDTree<data_t,traits> {bounds, nodes, primIDs, data, numPrims, numNodes}
//   buildTree(tree, d_points, N, BuildConfig{}, stream, memResource)
//   free(tree, stream, memResource)
//   stackBased::fcp(tree, q) / cct::fcp(tree, q)                (device)
//   stackBased::knn(list, tree, q) / cct::knn(list, tree, q)    (device)
//
// The build is deterministic: children are numbered in ascending parent
// order and degenerate splits alternate prims by rank in ascending prim id
// (the reference takes atomicAdd order for both, so its numbering is not
// reproducible run to run; the queries it answers are the same kind).
#pragma once

#include "cukd/knn.h"

namespace cukd {

struct BuildConfig {
  // make a leaf at or below this many prims; 0 means 8
  int makeLeafThreshold = 0;
};

template <typename data_t, typename data_traits = default_data_traits<data_t>>
struct SpatialKDTree {
  using point_t = typename data_traits::point_t;
  using scalar_t = typename point_traits<point_t>::scalar_t;
  using box_t = cukd::box_t<point_t>;

  struct Node {
    scalar_t pos;     // split position (inner nodes)
    uint32_t offset;  // first child (inner) or first primIDs slot (leaf)
    uint16_t count;   // prims in the leaf; 0 for inner nodes
    int16_t dim;      // split dim (inner nodes)
  };

  box_t bounds;
  Node *nodes = nullptr;
  uint32_t *primIDs = nullptr;
  const data_t *data = nullptr;
  int numPrims = 0;
  int numNodes = 0;
};

namespace detail {

// Caller-provided storage for the first entries of the stack-based
// traversal's stack: entry i at node[i * stride] / d2[i * stride] (for
// example one column per lane in shared memory, so the entries a kernel
// touches most are not in per-lane scratch).
struct SpatialStackCols {
  int *node;
  float *d2;
  int stride;
};

// Stack-based (CCT = false) or closest-corner (CCT = true) traversal of a
// spatial tree with the reference's 50-entry stack; returns false when
// that stack overflows (the reference prints "STACK OVERFLOW" and returns
// -1, fcp.h:317-320, knn.h:620-623).  With COLS > 0 (stack-based only),
// entries 0 .. COLS-1 live in 'cols' and the rest in the local array.
template <int D, bool CCT, int COLS = 0, class R, class NodeT, class Acc>
inline __both__ bool spatial_traverse(R &result, const float (&q)[D], const NodeT *nodes,
                                      const uint32_t *prim_ids, const float (&lo)[D],
                                      const float (&hi)[D], const Acc &data, unsigned &visits,
                                      SpatialStackCols cols = SpatialStackCols{nullptr, nullptr, 0})
{
  enum { DEPTH = 50, LOCAL = CCT ? DEPTH : DEPTH - COLS };
  static_assert(COLS >= 0 && COLS < DEPTH && (COLS == 0 || !CCT), "stack columns: stack-based only");
  float cull = result.initialCullDist2();
  int st_node[LOCAL];
  float st_d2[CCT ? 1 : LOCAL];
  float st_corner[CCT ? DEPTH : 1][D];
  int sp = 0;
  int node_id = 0;
  float closest[D];
  if (CCT) {
#pragma unroll
    for (int d = 0; d < D; d++) {
      float v = q[d];
      v = v < lo[d] ? lo[d] : v;
      v = hi[d] < v ? hi[d] : v;
      closest[d] = v;
    }
    if (sqr_dist<D>(q, closest) > cull) return true;
  }
  for (;;) {
    NodeT node;
    for (;;) {
      visits++;
      node = nodes[node_id];
      // an empty root leaf (no points) has count 0 and offset ~0u
      if (node.count || node.offset == 0xffffffffu) break;
      const int dim = node.dim;
      const float pos = (float)node.pos;
      const float qc = pick<D>(q, dim);
      const bool left_close = qc < pos;
      const int lchild = (int)node.offset, rchild = (int)node.offset + 1;
      const int far_c = left_close ? rchild : lchild;
      if (CCT) {
        float corner[D];
#pragma unroll
        for (int d = 0; d < D; d++) corner[d] = d == dim ? pos : closest[d];
        if (sqr_dist<D>(corner, q) < cull) {
#pragma unroll
          for (int d = 0; d < D; d++) st_corner[CCT ? sp : 0][d] = corner[d];
          st_node[sp] = far_c;
          if (++sp >= DEPTH) return false;
        }
      } else {
        const float diff = qc - pos;
        const float plane2 = diff * diff;
        if (plane2 < cull) {
          if (COLS > 0 && sp < COLS) {
            cols.node[sp * cols.stride] = far_c;
            cols.d2[sp * cols.stride] = plane2;
          } else {
            st_node[sp - COLS] = far_c;
            st_d2[CCT ? 0 : sp - COLS] = plane2;
          }
          if (++sp >= DEPTH) return false;
        }
      }
      node_id = left_close ? lchild : rchild;
    }
    for (int i = 0; i < node.count; i++) {
      const int prim = (int)prim_ids[node.offset + i];
      visits++;
      float p[D];
      data.template point<D>(prim, p);
      cull = result.processCandidate(prim, sqr_dist<D>(p, q));
    }
    for (;;) {
      if (sp == 0) return true;
      --sp;
      if (CCT) {
#pragma unroll
        for (int d = 0; d < D; d++) closest[d] = st_corner[CCT ? sp : 0][d];
        if (sqr_dist<D>(closest, q) >= cull) continue;
        node_id = st_node[sp];
        break;
      }
      const bool in_cols = COLS > 0 && sp < COLS;
      if ((in_cols ? cols.d2[sp * cols.stride] : st_d2[CCT ? 0 : sp - COLS]) >= cull) continue;
      node_id = in_cols ? cols.node[sp * cols.stride] : st_node[sp - COLS];
      break;
    }
  }
}

template <typename data_t, typename data_traits, bool CCT, class R>
inline __both__ bool spatial_query(R &result, const SpatialKDTree<data_t, data_traits> &tree,
                                   typename data_traits::point_t queryPoint)
{
  enum { D = point_traits<typename data_traits::point_t>::num_dims };
  float q[D], lo[D], hi[D];
  to_floats(queryPoint, q);
  to_floats(tree.bounds.lower, lo);
  to_floats(tree.bounds.upper, hi);
  unsigned visits = 0;
  const bool ok = spatial_traverse<D, CCT>(result, q, tree.nodes, tree.primIDs, lo, hi,
                                           TypedNodes<data_t, data_traits>{tree.data, tree.numPrims},
                                           visits);
  count_visits(visits);
  return ok;
}
}  // namespace detail

namespace stackBased {
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ int fcp(const SpatialKDTree<data_t, data_traits> &tree,
                        typename data_traits::point_t queryPoint,
                        FcpSearchParams params = FcpSearchParams{})
{
  FCPResult r;
  r.clear(sqr(params.cutOffRadius));
  if (!detail::spatial_query<data_t, data_traits, false>(r, tree, queryPoint)) return -1;
  return r.returnValue();
}
template <typename CandidateList, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ float knn(CandidateList &result, const SpatialKDTree<data_t, data_traits> &tree,
                          typename data_traits::point_t queryPoint)
{
  if (!detail::spatial_query<data_t, data_traits, false>(result, tree, queryPoint)) return -1.f;
  return result.returnValue();
}
}  // namespace stackBased

namespace cct {
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ int fcp(const SpatialKDTree<data_t, data_traits> &tree,
                        typename data_traits::point_t queryPoint,
                        FcpSearchParams params = FcpSearchParams{})
{
  FCPResult r;
  r.clear(sqr(params.cutOffRadius));
  if (!detail::spatial_query<data_t, data_traits, true>(r, tree, queryPoint)) return -1;
  return r.returnValue();
}
template <typename CandidateList, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ float knn(CandidateList &result, const SpatialKDTree<data_t, data_traits> &tree,
                          typename data_traits::point_t queryPoint)
{
  if (!detail::spatial_query<data_t, data_traits, true>(result, tree, queryPoint)) return -1.f;
  return result.returnValue();
}
}  // namespace cct

// ---------------------------------------------------------------------------
// host side: build / free (synchronise 'stream', as the reference does)
// ---------------------------------------------------------------------------
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void buildTree(SpatialKDTree<data_t, data_traits> &tree, data_t *d_points, int numPrims,
                      BuildConfig buildConfig = {}, hipStream_t stream = 0,
                      GpuMemoryResource &memResource = defaultGpuMemResource())
{
  using scalar_t = typename SpatialKDTree<data_t, data_traits>::scalar_t;
  const cukd_layout L = layout_of<data_t, data_traits>();
  cukd_mem_resource mr = detail::bridge(memResource);
  cukd_spatial_tree raw;
  detail::check(cukd_spatial_build(&raw, d_points, numPrims, &L, buildConfig.makeLeafThreshold,
                                   stream, &mr),
                "buildTree(SpatialKDTree)");
  enum { D = point_traits<typename data_traits::point_t>::num_dims };
  // float points: float bounds; int points: the int32 bounds
  for (int d = 0; d < D; d++) {
    const scalar_t lo = std::is_same<scalar_t, float>::value ? (scalar_t)raw.bounds[d]
                                                             : (scalar_t)raw.ibounds[d];
    const scalar_t hi = std::is_same<scalar_t, float>::value ? (scalar_t)raw.bounds[D + d]
                                                             : (scalar_t)raw.ibounds[D + d];
    point_traits<typename data_traits::point_t>::set_coord(tree.bounds.lower, d, lo);
    point_traits<typename data_traits::point_t>::set_coord(tree.bounds.upper, d, hi);
  }
  tree.nodes = (typename SpatialKDTree<data_t, data_traits>::Node *)raw.nodes;
  tree.primIDs = raw.prim_ids;
  tree.data = d_points;
  tree.numPrims = raw.num_prims;
  tree.numNodes = raw.num_nodes;
}

template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void free(SpatialKDTree<data_t, data_traits> &tree, hipStream_t stream = 0,
                 GpuMemoryResource &memResource = defaultGpuMemResource())
{
  cukd_spatial_tree raw;
  raw.nodes = (cukd_spatial_node *)tree.nodes;
  raw.prim_ids = tree.primIDs;
  cukd_mem_resource mr = detail::bridge(memResource);
  detail::check(cukd_spatial_free(&raw, stream, &mr), "free(SpatialKDTree)");
  tree.nodes = nullptr;
  tree.primIDs = nullptr;
  tree.numNodes = 0;
  tree.numPrims = 0;
}

}  // namespace cukd
