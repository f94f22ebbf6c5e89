// cukd/data.h -- the data_traits concept (drop-in for the reference's
// cukd/data.h): how to get a point, a coordinate and an optional split dim
// out of a user's data_t.
#pragma once

#include "cukd/box.h"
#include "cukd/cukd-math.h"

namespace cukd {

// data_t == point_t, round-robin split dims, no payload.
template <typename _point_t, typename _point_traits = ::cukd::point_traits<_point_t>>
struct default_data_traits {
  using point_t = _point_t;
  using point_traits = _point_traits;
  using data_t = _point_t;
  using scalar_t = typename point_traits::scalar_t;

  static inline __both__ const point_t &get_point(const data_t &n) { return n; }
  static inline __both__ scalar_t get_coord(const data_t &n, int d)
  {
    return point_traits::get_coord(get_point(n), d);
  }

  enum { has_explicit_dim = false };
  static inline __both__ int get_dim(const data_t &) { return -1; }
  static inline __both__ void set_dim(data_t &, int) {}
};

}  // namespace cukd
