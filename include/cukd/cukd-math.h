// cukd/cukd-math.h -- point types, point_traits and distance math.
//
// Drop-in for cudaKDTree's cukd/cukd-math.h (v1.0.1): the same names
// (scalar_type_of, num_dims_of, get_coord/set_coord, point_traits,
// vec_float<N>, sqrDistance, distance) over HIP's vector types.
// Distances are evaluated as d = a - b per component and
// dist2 = ((d0*d0 + d1*d1) + d2*d2) + ..., never contracted to FMA, so
// device and host agree bit for bit.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "cukd/common.h"

namespace cukd {

template <typename T> struct scalar_type_of;
template <> struct scalar_type_of<float2> { using type = float; };
template <> struct scalar_type_of<float3> { using type = float; };
template <> struct scalar_type_of<float4> { using type = float; };
template <> struct scalar_type_of<int2> { using type = int; };
template <> struct scalar_type_of<int3> { using type = int; };
template <> struct scalar_type_of<int4> { using type = int; };

template <typename T> struct num_dims_of;
template <> struct num_dims_of<float2> { enum { value = 2 }; };
template <> struct num_dims_of<float3> { enum { value = 3 }; };
template <> struct num_dims_of<float4> { enum { value = 4 }; };
template <> struct num_dims_of<int2> { enum { value = 2 }; };
template <> struct num_dims_of<int3> { enum { value = 3 }; };
template <> struct num_dims_of<int4> { enum { value = 4 }; };

// coordinate access on HIP vector types: member d of (x, y, z, w)
template <typename V>
inline __both__ auto vec_coord(const V &v, int d) -> decltype(v.x)
{
  const decltype(v.x) *c = &v.x;
  return c[d];
}

inline __both__ float get_coord(const float2 &v, int d) { return vec_coord(v, d); }
inline __both__ float get_coord(const float3 &v, int d) { return vec_coord(v, d); }
inline __both__ float get_coord(const float4 &v, int d) { return vec_coord(v, d); }
inline __both__ int get_coord(const int2 &v, int d) { return vec_coord(v, d); }
inline __both__ int get_coord(const int3 &v, int d) { return vec_coord(v, d); }
inline __both__ int get_coord(const int4 &v, int d) { return vec_coord(v, d); }

template <typename V, typename S>
inline __both__ void vec_set_coord(V &v, int d, S s)
{
  decltype(v.x) *c = &v.x;
  c[d] = s;
}
inline __both__ void set_coord(float2 &v, int d, float s) { vec_set_coord(v, d, s); }
inline __both__ void set_coord(float3 &v, int d, float s) { vec_set_coord(v, d, s); }
inline __both__ void set_coord(float4 &v, int d, float s) { vec_set_coord(v, d, s); }
inline __both__ void set_coord(int2 &v, int d, int s) { vec_set_coord(v, d, s); }
inline __both__ void set_coord(int3 &v, int d, int s) { vec_set_coord(v, d, s); }
inline __both__ void set_coord(int4 &v, int d, int s) { vec_set_coord(v, d, s); }

// arbitrary-dimension float point (cukd-math.h:135-181 in the reference)
template <int N>
struct vec_float {
  float v[N];
};
template <int N> struct scalar_type_of<vec_float<N>> { using type = float; };
template <int N> struct num_dims_of<vec_float<N>> { enum { value = N }; };
template <int N>
inline __both__ float get_coord(const vec_float<N> &p, int d) { return p.v[d]; }
template <int N>
inline __both__ void set_coord(vec_float<N> &p, int d, float s) { p.v[d] = s; }

// point_traits: scalar type, dimension count, coordinate get/set
template <typename point_t>
struct point_traits {
  enum { num_dims = num_dims_of<point_t>::value };
  using scalar_t = typename scalar_type_of<point_t>::type;
  static inline __both__ scalar_t get_coord(const point_t &p, int d) { return ::cukd::get_coord(p, d); }
  static inline __both__ void set_coord(point_t &p, int d, scalar_t s) { ::cukd::set_coord(p, d, s); }
};

template <typename scalar_t>
inline __both__ scalar_t sqr(scalar_t f) { return f * f; }

// dist2 with the pinned evaluation order
template <typename point_t>
inline __both__ float sqrDistance(const point_t &a, const point_t &b)
{
  using pt = point_traits<point_t>;
  float df = (float)pt::get_coord(a, 0) - (float)pt::get_coord(b, 0);
  float s = df * df;
  for (int i = 1; i < pt::num_dims; i++) {
    df = (float)pt::get_coord(a, i) - (float)pt::get_coord(b, i);
    s = s + df * df;
  }
  return s;
}

template <typename point_t>
inline __both__ float distance(const point_t &a, const point_t &b)
{
  return sqrtf(sqrDistance(a, b));
}

// component-wise min / max used by box_t
template <typename point_t>
inline __both__ point_t min_point(const point_t &a, const point_t &b)
{
  using pt = point_traits<point_t>;
  point_t r = a;
  for (int d = 0; d < pt::num_dims; d++)
    pt::set_coord(r, d, pt::get_coord(b, d) < pt::get_coord(a, d) ? pt::get_coord(b, d)
                                                                   : pt::get_coord(a, d));
  return r;
}
template <typename point_t>
inline __both__ point_t max_point(const point_t &a, const point_t &b)
{
  using pt = point_traits<point_t>;
  point_t r = a;
  for (int d = 0; d < pt::num_dims; d++)
    pt::set_coord(r, d, pt::get_coord(a, d) < pt::get_coord(b, d) ? pt::get_coord(b, d)
                                                                   : pt::get_coord(a, d));
  return r;
}

}  // namespace cukd
