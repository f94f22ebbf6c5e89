// cukd/box.h -- axis-aligned box over a point type (drop-in for the
// reference's cukd/box.h: box_t, widestDimension, project, sqrDistance).
#pragma once

#include <limits.h>
#include <math.h>

#include "cukd/cukd-math.h"

namespace cukd {

template <typename T> inline __both__ T empty_box_lower();
template <typename T> inline __both__ T empty_box_upper();
template <> inline __both__ float empty_box_lower<float>() { return INFINITY; }
template <> inline __both__ float empty_box_upper<float>() { return -INFINITY; }
template <> inline __both__ int empty_box_lower<int>() { return INT_MAX; }
template <> inline __both__ int empty_box_upper<int>() { return INT_MIN; }

// Memory layout: lower then upper point; the C ABI's box buffers
// (2 * num_dims scalars) have the same layout.
template <typename point_t>
struct box_t {
  using point_traits = ::cukd::point_traits<point_t>;
  using scalar_t = typename point_traits::scalar_t;

  // The dim with the widest extent; equal extents pick the later dim and
  // a box with no positive extent reports dim 0 unless a later dim ties
  // at zero (same rule as the reference, box.h:110-125 there).
  inline __both__ int widestDimension() const
  {
    int best = 0;
    scalar_t wb = scalar_t(0);
    for (int d = 0; d < point_traits::num_dims; d++) {
      const scalar_t w = point_traits::get_coord(upper, d) - point_traits::get_coord(lower, d);
      if (w < wb) continue;
      wb = w;
      best = d;
    }
    return best;
  }

  inline __both__ bool contains(const point_t &p) const
  {
    for (int d = 0; d < point_traits::num_dims; d++) {
      if (point_traits::get_coord(p, d) < point_traits::get_coord(lower, d)) return false;
      if (point_traits::get_coord(p, d) > point_traits::get_coord(upper, d)) return false;
    }
    return true;
  }

  inline __both__ void grow(const point_t &p)
  {
    lower = min_point(lower, p);
    upper = max_point(upper, p);
  }

  inline __both__ void setEmpty()
  {
    for (int d = 0; d < point_traits::num_dims; d++) {
      point_traits::set_coord(lower, d, empty_box_lower<scalar_t>());
      point_traits::set_coord(upper, d, empty_box_upper<scalar_t>());
    }
  }

  inline __both__ void setInfinite()
  {
    for (int d = 0; d < point_traits::num_dims; d++) {
      point_traits::set_coord(lower, d, empty_box_upper<scalar_t>());
      point_traits::set_coord(upper, d, empty_box_lower<scalar_t>());
    }
  }

  point_t lower, upper;
};

// closest point of the box to 'point'
template <typename point_t>
inline __both__ point_t project(const box_t<point_t> &box, const point_t &point)
{
  return min_point(max_point(point, box.lower), box.upper);
}

template <typename point_t>
inline __both__ float sqrDistance(const box_t<point_t> &box, const point_t &point)
{
  return sqrDistance(project(box, point), point);
}

}  // namespace cukd
