// cukd/traverse.h -- the three traversals of a left-balanced k-d tree.
//
// One implementation serves both the device-inline API (user kernels call
// traverse_default / traverse_stack_free / traverse_cct, or the fcp/knn
// wrappers, exactly as with the reference) and the library's batched
// kernels.  The core works on a node accessor:
//     acc.size()                 number of nodes
//     acc.template point<D>(i,p) coordinates of node i as float[D]
//     acc.dim(i)                 split dim of node i
// and on any result type with initialCullDist2() / processCandidate().
// Control flow follows the reference line for line in effect:
//   stack-based  traverse-default-stack-based.h:24-108
//   stack-free   traverse-stack-free.h:24-117
//   cct          traverse-cct.h:25-99
#pragma once
#pragma clang fp contract(off)

#include "cukd/box.h"
#include "cukd/data.h"
#include "cukd/helpers.h"

namespace cukd {
namespace detail {

// p[d] for a run-time d, as a blend of the D registers (a plain p[d]
// or a select chain on floats gets turned into an indexed load, which
// spills p to LDS on every visit)
template <int D>
inline __both__ float pick(const float (&p)[D], int d)
{
#ifdef __HIP_DEVICE_COMPILE__
  unsigned v = __float_as_uint(p[0]);
#pragma unroll
  for (int i = 1; i < D; i++) {
    const unsigned m = 0u - (unsigned)(d == i);
    v = (__float_as_uint(p[i]) & m) | (v & ~m);
  }
  return __uint_as_float(v);
#else
  return p[d];
#endif
}

template <int D>
inline __both__ float sqr_dist(const float (&a)[D], const float (&b)[D])
{
  float df = a[0] - b[0];
  float s = df * df;
#pragma unroll
  for (int i = 1; i < D; i++) {
    df = a[i] - b[i];
    s = s + df * df;
  }
  return s;
}

template <int D>
inline __both__ int round_robin_dim(int node)
{
  const int l = BinaryTree::levelOf(node);
  return (D & (D - 1)) == 0 ? (l & (D - 1)) : l % D;
}

// Closest-point-first DFS (the reference keeps a 30-entry stack of
// (far child, plane dist2)).  The stack is always ordered by depth -- every
// push while descending is deeper than the previous one, and after popping
// depth d only deeper entries are pushed -- so it is a set of depths: bit d
// of 'pending' marks that the far child at depth d, the sibling of the
// current path's node at that depth, is still to be visited.  Its plane
// distance is recomputed from its parent on pop (the same float operations
// as at push time), so the visit order and results equal the reference's.
template <int D, class Acc, class R>
inline __both__ void traverse_stack_based(R &result, const float (&q)[D], const Acc &acc,
                                          unsigned &visits)
{
  const int n = acc.size();
  float cull = result.initialCullDist2();
  unsigned pending = 0;
  int curr = 0;
  for (;;) {
    while (curr < n) {
      float p[D];
      acc.template point<D>(curr, p);
      visits++;
      const int dim = acc.dim(curr);
      cull = result.processCandidate(curr, sqr_dist<D>(p, q));
      const float nc = pick<D>(p, dim);
      const float qc = pick<D>(q, dim);
      const bool left_close = qc < nc;
      const int l = 2 * curr + 1;
      const int far_c = left_close ? l + 1 : l;
      const float diff = qc - nc;
      if (diff * diff < cull && far_c < n) pending |= 1u << (BinaryTree::levelOf(curr) + 1);
      curr = left_close ? l : l + 1;
    }
    for (;;) {
      if (!pending) return;
#ifdef __HIP_DEVICE_COMPILE__
      const int d = 31 - __clz((int)pending);
#else
      const int d = 31 - __builtin_clz(pending);
#endif
      pending &= ~(1u << d);
      const int lc = BinaryTree::levelOf(curr);
      const int on_path = ((curr + 1) >> (lc - d)) - 1;
      const int far_c = (on_path & 1) ? on_path + 1 : on_path - 1;
      const int parent = (far_c - 1) >> 1;
      float p[D];
      acc.template point<D>(parent, p);
      const int dim = acc.dim(parent);
      const float diff = pick<D>(q, dim) - pick<D>(p, dim);
      if (diff * diff >= cull) continue;
      curr = far_c;
      break;
    }
  }
}

// parent/child state machine, no stack
template <int D, class Acc, class R>
inline __both__ void traverse_stack_free(R &result, const float (&q)[D], const Acc &acc,
                                         float eps, unsigned &visits)
{
  const int n = acc.size();
  if (n <= 0) return;
  const float eps_err = 1.f + eps;
  float cull = result.initialCullDist2();
  int prev = -1, curr = 0;
  for (;;) {
    const int parent = (curr + 1) / 2 - 1;
    if (curr >= n) {
      // a missing close child: pretend it was done
      prev = curr;
      curr = parent;
      continue;
    }
    float p[D];
    acc.template point<D>(curr, p);
    visits++;
    const int child = 2 * curr + 1;
    if (!(prev >= child)) cull = result.processCandidate(curr, sqr_dist<D>(q, p));
    const int dim = acc.dim(curr);
    const float dd = pick<D>(q, dim) - pick<D>(p, dim);
    const int side = dd > 0.f;
    const int close_c = 2 * curr + 1 + side;
    const int far_c = 2 * curr + 2 - side;
    int next;
    if (prev == close_c)
      next = (far_c < n && (dd * dd) * eps_err < cull) ? far_c : parent;
    else if (prev == far_c)
      next = parent;
    else
      next = child < n ? close_c : parent;
    if (next == -1) return;
    prev = curr;
    curr = next;
  }
}

// closest-corner tracking: each stack entry carries the closest point of
// the far subtree's domain
template <int D, class Acc, class R>
inline __both__ void traverse_cct(R &result, const float (&q)[D], const float (&lo)[D],
                                  const float (&hi)[D], const Acc &acc, unsigned &visits)
{
  const int n = acc.size();
  float cull = result.initialCullDist2();
  int st_node[30];
  float st_corner[30][D];
  int sp = 0;
  int node = 0;
  float closest[D];
#pragma unroll
  for (int d = 0; d < D; d++) {
    float v = q[d];
    v = v < lo[d] ? lo[d] : v;
    v = hi[d] < v ? hi[d] : v;
    closest[d] = v;
  }
  if (sqr_dist<D>(q, closest) > cull) return;
  for (;;) {
    if (node >= n) {
      for (;;) {
        if (sp == 0) return;
        --sp;
#pragma unroll
        for (int d = 0; d < D; d++) closest[d] = st_corner[sp][d];
        if (sqr_dist<D>(closest, q) >= cull) continue;
        node = st_node[sp];
        break;
      }
    }
    float p[D];
    acc.template point<D>(node, p);
    visits++;
    cull = result.processCandidate(node, sqr_dist<D>(p, q));
    const int dim = acc.dim(node);
    const float nc = pick<D>(p, dim);
    const bool left_close = pick<D>(q, dim) < nc;
    const int l = 2 * node + 1;
    const int far_c = left_close ? l + 1 : l;
    float corner[D];
#pragma unroll
    for (int d = 0; d < D; d++) corner[d] = d == dim ? nc : closest[d];
    if (far_c < n && sqr_dist<D>(corner, q) < cull) {
#pragma unroll
      for (int d = 0; d < D; d++) st_corner[sp][d] = corner[d];
      st_node[sp] = far_c;
      sp++;
    }
    node = left_close ? l : l + 1;
  }
}

// accessor over a typed data_t array described by data_traits
template <typename data_t, typename data_traits>
struct TypedNodes {
  const data_t *nodes;
  int n;
  enum { D = point_traits<typename data_traits::point_t>::num_dims };
  inline __both__ int size() const { return n; }
  template <int DD>
  inline __both__ void point(int i, float (&p)[DD]) const
  {
    for (int d = 0; d < DD; d++) p[d] = (float)data_traits::get_coord(nodes[i], d);
  }
  inline __both__ int dim(int i) const
  {
    if (data_traits::has_explicit_dim) return data_traits::get_dim(nodes[i]);
    return round_robin_dim<D>(i);
  }
};

template <typename point_t, int D>
inline __both__ void to_floats(const point_t &p, float (&q)[D])
{
  for (int d = 0; d < D; d++) q[d] = (float)point_traits<point_t>::get_coord(p, d);
}

}  // namespace detail

// ---------------------------------------------------------------------------
// the reference's traversal entry points (typed)
// ---------------------------------------------------------------------------

template <typename result_t, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ void traverse_default(result_t &result, typename data_traits::point_t queryPoint,
                                      const data_t *d_nodes, int numPoints)
{
  enum { D = point_traits<typename data_traits::point_t>::num_dims };
  float q[D];
  detail::to_floats(queryPoint, q);
  unsigned visits = 0;
  detail::traverse_stack_based<D>(result, q, detail::TypedNodes<data_t, data_traits>{d_nodes, numPoints},
                                  visits);
  detail::count_visits(visits);
}

template <typename result_t, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ void traverse_stack_free(result_t &result, typename data_traits::point_t queryPoint,
                                         const data_t *d_nodes, int N, float eps = 0.0f)
{
  enum { D = point_traits<typename data_traits::point_t>::num_dims };
  float q[D];
  detail::to_floats(queryPoint, q);
  unsigned visits = 0;
  detail::traverse_stack_free<D>(result, q, detail::TypedNodes<data_t, data_traits>{d_nodes, N}, eps,
                                 visits);
  detail::count_visits(visits);
}

template <typename result_t, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ void traverse_cct(result_t &result, typename data_traits::point_t queryPoint,
                                  const box_t<typename data_traits::point_t> d_bounds,
                                  const data_t *d_nodes, int numPoints)
{
  enum { D = point_traits<typename data_traits::point_t>::num_dims };
  float q[D], lo[D], hi[D];
  detail::to_floats(queryPoint, q);
  detail::to_floats(d_bounds.lower, lo);
  detail::to_floats(d_bounds.upper, hi);
  unsigned visits = 0;
  detail::traverse_cct<D>(result, q, lo, hi, detail::TypedNodes<data_t, data_traits>{d_nodes, numPoints},
                          visits);
  detail::count_visits(visits);
}

}  // namespace cukd
