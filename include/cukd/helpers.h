// cukd/helpers.h -- memory resources, tree index math, and the mapping of a
// data_traits description onto the C ABI's cukd_layout.
#pragma once

#include <stdexcept>
#include <string.h>
#include <string>
#include <type_traits>

#include "cukd/data.h"
#include "cukd_amd.h"

namespace cukd {

// ---------------------------------------------------------------------------
// memory resources (the reference's GpuMemoryResource, helpers.h:31-71)
// ---------------------------------------------------------------------------
struct GpuMemoryResource {
  virtual hipError_t malloc(void **ptr, size_t size, hipStream_t s) = 0;
  virtual hipError_t free(void *ptr, hipStream_t s) = 0;
  virtual ~GpuMemoryResource() {}
};

// stream-ordered allocation (the default)
struct AsyncGpuMemoryResource final : GpuMemoryResource {
  hipError_t malloc(void **ptr, size_t size, hipStream_t s) override { return hipMallocAsync(ptr, size, s); }
  hipError_t free(void *ptr, hipStream_t s) override { return hipFreeAsync(ptr, s); }
};

struct ManagedMemMemoryResource : GpuMemoryResource {
  hipError_t malloc(void **ptr, size_t size, hipStream_t s) override
  {
    const hipError_t e = hipStreamSynchronize(s);
    return e != hipSuccess ? e : hipMallocManaged(ptr, size);
  }
  hipError_t free(void *ptr, hipStream_t s) override
  {
    const hipError_t e = hipStreamSynchronize(s);
    return e != hipSuccess ? e : hipFree(ptr);
  }
};

inline GpuMemoryResource &defaultGpuMemResource()
{
  static AsyncGpuMemoryResource r;
  return r;
}

namespace detail {
inline int mem_malloc_tramp(void *ctx, void **ptr, size_t bytes, hipStream_t s)
{
  return ((GpuMemoryResource *)ctx)->malloc(ptr, bytes, s) == hipSuccess ? 0 : 1;
}
inline int mem_free_tramp(void *ctx, void *ptr, hipStream_t s)
{
  return ((GpuMemoryResource *)ctx)->free(ptr, s) == hipSuccess ? 0 : 1;
}
inline cukd_mem_resource bridge(GpuMemoryResource &m)
{
  cukd_mem_resource r;
  r.ctx = &m;
  r.malloc_fn = mem_malloc_tramp;
  r.free_fn = mem_free_tramp;
  return r;
}
inline void check(int rc, const char *what)
{
  if (rc != CUKD_OK)
    throw std::runtime_error(std::string("cukd::") + what + ": " + cukd_last_error());
}
}  // namespace detail

// ---------------------------------------------------------------------------
// implicit binary tree index math (the reference's BinaryTree etc.)
// ---------------------------------------------------------------------------
struct BinaryTree {
  static inline __both__ int rootNode() { return 0; }
  static inline __both__ int parentOf(int n) { return (n - 1) / 2; }
  static inline __both__ int leftChildOf(int n) { return 2 * n + 1; }
  static inline __both__ int rightChildOf(int n) { return 2 * n + 2; }
  static inline __both__ int firstNodeInLevel(int L) { return (1 << L) - 1; }
  static inline __both__ int levelOf(int n)
  {
#ifdef __HIP_DEVICE_COMPILE__
    return 31 - __clz(n + 1);
#else
    return 31 - __builtin_clz((unsigned)(n + 1));
#endif
  }
  static inline __both__ int numLevelsFor(int numPoints) { return levelOf(numPoints - 1) + 1; }
};

// ---------------------------------------------------------------------------
// data_traits -> cukd_layout
// ---------------------------------------------------------------------------

// Derives the runtime layout of data_t by probing data_traits on a zeroed
// object: the point must be a sub-object of data_t with contiguous 4-byte
// coordinates, and set_dim (when has_explicit_dim) must write one 1-, 2- or
// 4-byte field.  Throws std::runtime_error otherwise.
template <typename data_t, typename data_traits>
inline cukd_layout layout_of()
{
  using point_t = typename data_traits::point_t;
  using pt = point_traits<point_t>;
  using scalar_t = typename pt::scalar_t;
  static_assert(std::is_same<scalar_t, float>::value || std::is_same<scalar_t, int>::value,
                "cukd: coordinates must be float or int");
  static_assert(sizeof(data_t) % 4 == 0, "cukd: sizeof(data_t) must be a multiple of 4");
  cukd_layout L;
  L.num_dims = pt::num_dims;
  L.scalar_type = std::is_same<scalar_t, float>::value ? CUKD_SCALAR_F32 : CUKD_SCALAR_I32;
  L.stride = (int)sizeof(data_t);
  L.dim_offset = -1;
  L.dim_bytes = 4;
  alignas(16) unsigned char raw[sizeof(data_t)];
  memset(raw, 0, sizeof(raw));
  data_t &probe = *reinterpret_cast<data_t *>(raw);
  const point_t &p = data_traits::get_point(probe);
  const long off = (const unsigned char *)&p - raw;
  if (off < 0 || off + (long)(4 * pt::num_dims) > (long)sizeof(data_t))
    throw std::runtime_error("cukd: data_traits::get_point must return a reference into data_t");
  L.point_offset = (int)off;
  for (int d = 0; d < pt::num_dims; d++) {
    unsigned char before[sizeof(data_t)];
    memcpy(before, raw, sizeof(raw));
    pt::set_coord(const_cast<point_t &>(p), d, (scalar_t)1);
    if (memcmp(before, raw, sizeof(raw)) == 0 || (raw[off + 4 * d] == 0 && raw[off + 4 * d + 3] == 0))
      throw std::runtime_error("cukd: point coordinates must be contiguous 4-byte scalars");
  }
  if (data_traits::has_explicit_dim) {
    memset(raw, 0, sizeof(raw));
    data_traits::set_dim(probe, 0x01010101);
    int first = -1, count = 0;
    for (int i = 0; i < (int)sizeof(data_t); i++)
      if (raw[i]) { if (first < 0) first = i; count++; }
    if (first < 0 || (count != 1 && count != 2 && count != 4))
      throw std::runtime_error("cukd: data_traits::set_dim must write a 1, 2 or 4 byte field");
    L.dim_offset = first;
    L.dim_bytes = count;
  }
  return L;
}

}  // namespace cukd
