// capi.hip -- the extern "C" boundary (include/cukd_amd.h): argument checks,
// error reporting, the default memory resource, and dispatch to the
// builder and query kernels.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <map>
#include <mutex>

#include "cukd_common.h"
#include "query_impl.h"
#include "forest_plan.h"

namespace cukd_amd {

static thread_local char g_last_error[512] = "";
static unsigned long long *g_stats = nullptr;

void set_last_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

int MemResource::alloc(void **p, size_t bytes, hipStream_t s) const
{
  *p = nullptr;
  if (bytes == 0) bytes = 256;
  if (user && user->malloc_fn) {
    if (user->malloc_fn(user->ctx, p, bytes, s) != 0 || !*p) {
      set_last_error("memory resource failed to allocate %zu bytes", bytes);
      return -1;
    }
    return 0;
  }
  const hipError_t e = hipMallocAsync(p, bytes, s);
  if (e != hipSuccess || !*p) {
    set_last_error("hipMallocAsync(%zu) failed: %s", bytes, hipGetErrorString(e));
    return -1;
  }
  return 0;
}

void MemResource::release(void *p, hipStream_t s) const
{
  if (!p) return;
  if (user && user->free_fn) {
    user->free_fn(user->ctx, p, s);
    return;
  }
  (void)hipFreeAsync(p, s);
}

int build_tree_tl(void *d_points, int n, const cukd_layout &L, void *d_bounds, hipStream_t s,
                  const MemResource &mem, bool carry_rank);
size_t build_tl_temp_bytes(int n, const cukd_layout &L, bool carry_rank);
int build_tree_impl(void *d_points, int n, const cukd_layout &L, void *d_bounds,
                    hipStream_t s, const MemResource &mem);
size_t build_temp_bytes(int n, const cukd_layout &L);
int build_bounds_only(void *d_bounds, const void *d_points, int n, const cukd_layout &L,
                      hipStream_t s);
int screen_nan(const void *d_points, int n, const cukd_layout &L, hipStream_t stream);

size_t query_order_temp_bytes(int nq);
int32_t *make_query_order(void *tmp, const float *q, int nq, int D, hipStream_t s);

hipError_t launch_query_1(int, const QueryArgs &, hipStream_t);
hipError_t launch_query_2(int, const QueryArgs &, hipStream_t);
hipError_t launch_query_3(int, const QueryArgs &, hipStream_t);
hipError_t launch_query_4(int, const QueryArgs &, hipStream_t);
hipError_t launch_query_5(int, const QueryArgs &, hipStream_t);
hipError_t launch_query_6(int, const QueryArgs &, hipStream_t);
hipError_t launch_query_7(int, const QueryArgs &, hipStream_t);
hipError_t launch_query_8(int, const QueryArgs &, hipStream_t);

hipError_t launch_query(int D, int method, const QueryArgs &a, hipStream_t s)
{
  switch (D) {
    case 1: return launch_query_1(method, a, s);
    case 2: return launch_query_2(method, a, s);
    case 3: return launch_query_3(method, a, s);
    case 4: return launch_query_4(method, a, s);
    case 5: return launch_query_5(method, a, s);
    case 6: return launch_query_6(method, a, s);
    case 7: return launch_query_7(method, a, s);
    default: return launch_query_8(method, a, s);
  }
}

static int check_layout(const cukd_layout *L)
{
  if (!L) { set_last_error("layout is null"); return CUKD_ERR_INVALID; }
  if (L->num_dims < 1 || L->num_dims > CUKD_MAX_DIMS) {
    set_last_error("num_dims %d not in 1..%d", L->num_dims, CUKD_MAX_DIMS);
    return CUKD_ERR_INVALID;
  }
  if (L->scalar_type != CUKD_SCALAR_F32 && L->scalar_type != CUKD_SCALAR_I32) {
    set_last_error("unknown scalar_type %d", L->scalar_type);
    return CUKD_ERR_INVALID;
  }
  if (L->stride < 4 || (L->stride & 3) || L->stride > 256) {
    set_last_error("stride %d must be a multiple of 4 in [4,256]", L->stride);
    return CUKD_ERR_INVALID;
  }
  if (L->point_offset < 0 || (L->point_offset & 3) ||
      L->point_offset + 4 * L->num_dims > L->stride) {
    set_last_error("point_offset %d does not fit %d coordinates in stride %d",
                   L->point_offset, L->num_dims, L->stride);
    return CUKD_ERR_INVALID;
  }
  if (L->dim_offset >= 0) {
    if (L->dim_bytes != 1 && L->dim_bytes != 2 && L->dim_bytes != 4) {
      set_last_error("dim_bytes %d must be 1, 2 or 4", L->dim_bytes);
      return CUKD_ERR_INVALID;
    }
    if (L->dim_offset % L->dim_bytes || L->dim_offset + L->dim_bytes > L->stride) {
      set_last_error("dim field at %d (%d bytes) misaligned or outside the element",
                     L->dim_offset, L->dim_bytes);
      return CUKD_ERR_INVALID;
    }
    const int dlo = L->dim_offset, dhi = L->dim_offset + L->dim_bytes;
    const int plo = L->point_offset, phi = L->point_offset + 4 * L->num_dims;
    if (dlo < phi && plo < dhi) {
      set_last_error("dim field overlaps the coordinates");
      return CUKD_ERR_INVALID;
    }
  }
  return CUKD_OK;
}

}  // namespace cukd_amd

using namespace cukd_amd;

// opaque forest handle of the C ABI
struct cukd_forest {
  ForestPlan plan;
  uint8_t *ws = nullptr;        // build workspace, plan tables, tree offsets
  int32_t *tree_off = nullptr;  // num_trees + 1 element offsets (device)
  size_t bytes = 0;
  cukd_mem_resource mem{};
  bool has_mem = false;
  void *order_ws = nullptr;     // Morton-ordering scratch (cukd_forest_reserve_queries)
  size_t order_ws_bytes = 0;
  hipStream_t order_stream = nullptr;  // the only stream order_ws is used on
  std::mutex order_mu;                 // guards order_ws / order_stream
};

namespace cukd_amd {
void clear_last_error() { g_last_error[0] = 0; }
int forest_num_dims(const cukd_forest *f) { return f->plan.L.num_dims; }
int forest_num_trees(const cukd_forest *f) { return f->plan.ntrees; }
}  // namespace cukd_amd

extern "C" {

const char *cukd_last_error(void) { return g_last_error; }

int cukd_version(void) { return 100; }

int cukd_set_traversal_stats(unsigned long long *d_counter)
{
  g_stats = d_counter;
  return CUKD_OK;
}

int cukd_build_tree_temp_bytes(int32_t n, const cukd_layout *layout, int builder,
                               size_t *bytes)
{
  g_last_error[0] = 0;
  if (int rc = check_layout(layout)) return rc;
  if (!bytes) { set_last_error("bytes is null"); return CUKD_ERR_INVALID; }
  if (builder < CUKD_BUILDER_ID_DEFAULT || builder > CUKD_BUILDER_ID_INPLACE) {
    set_last_error("unknown builder %d", builder);
    return CUKD_ERR_INVALID;
  }
  if (n < 1) *bytes = 0;
  else if (builder == CUKD_BUILDER_ID_BITONIC) *bytes = build_tl_temp_bytes(n, *layout, true);
  else if (builder == CUKD_BUILDER_ID_INPLACE) *bytes = 0;
  else *bytes = build_temp_bytes(n, *layout);
  return CUKD_OK;
}

int cukd_build_tree(void *d_points, int32_t n, const cukd_layout *layout, int builder,
                    void *d_world_bounds, hipStream_t stream, const cukd_mem_resource *mem)
{
  g_last_error[0] = 0;
  if (int rc = check_layout(layout)) return rc;
  if (builder < CUKD_BUILDER_ID_DEFAULT || builder > CUKD_BUILDER_ID_INPLACE) {
    set_last_error("unknown builder %d", builder);
    return CUKD_ERR_INVALID;
  }
  if (n < 1) return CUKD_OK;
  if (n >= (1 << 30)) {
    set_last_error("n = %d exceeds the 2^30 limit of the 30-entry traversal stack", n);
    return CUKD_ERR_INVALID;
  }
  if (!d_points) { set_last_error("d_points is null"); return CUKD_ERR_INVALID; }
  if (layout->dim_offset >= 0 && !d_world_bounds) {
    set_last_error("asked to build k-d tree over nodes with explicit dims, "
                   "but no memory for world bounds provided");
    return CUKD_ERR_NO_BOUNDS;
  }
  if (int rc = screen_nan(d_points, n, *layout, stream)) return rc;
  MemResource mr{mem};
  if (builder == CUKD_BUILDER_ID_BITONIC || builder == CUKD_BUILDER_ID_INPLACE)
    return build_tree_tl(d_points, n, *layout, d_world_bounds, stream, mr,
                         builder == CUKD_BUILDER_ID_BITONIC);
  return build_tree_impl(d_points, n, *layout, d_world_bounds, stream, mr);
}

int cukd_build_tree_host(void *h_points, int32_t n, const cukd_layout *layout,
                         void *h_world_bounds)
{
  g_last_error[0] = 0;
  if (int rc = check_layout(layout)) return rc;
  if (n < 1) return CUKD_OK;
  if (!h_points) { set_last_error("points is null"); return CUKD_ERR_INVALID; }
  if (layout->dim_offset >= 0 && !h_world_bounds) {
    set_last_error("cukd::builder_host: asked to build k-d tree over nodes with explicit "
                   "dims, but no memory for world bounds provided");
    return CUKD_ERR_NO_BOUNDS;
  }
  const size_t bytes = (size_t)n * (size_t)layout->stride;
  const size_t box_bytes = (size_t)2 * layout->num_dims * 4;
  uint8_t *d = nullptr;
  hipError_t e = hipMalloc((void **)&d, bytes + 256 + box_bytes);
  if (e != hipSuccess) {
    set_last_error("hipMalloc(%zu) failed: %s", bytes + 256 + box_bytes, hipGetErrorString(e));
    return CUKD_ERR_ALLOC;
  }
  void *d_box = h_world_bounds ? (void *)(d + ((bytes + 255) & ~(size_t)255)) : nullptr;
  int rc = CUKD_OK;
  e = hipMemcpy(d, h_points, bytes, hipMemcpyDefault);
  if (e == hipSuccess) {
    rc = cukd_build_tree(d, n, layout, CUKD_BUILDER_ID_DEFAULT, d_box, nullptr, nullptr);
    if (rc == CUKD_OK) {
      e = hipMemcpy(h_points, d, bytes, hipMemcpyDefault);
      if (e == hipSuccess && d_box) e = hipMemcpy(h_world_bounds, d_box, box_bytes, hipMemcpyDefault);
    }
  }
  (void)hipFree(d);
  if (rc != CUKD_OK) return rc;
  if (e != hipSuccess) {
    set_last_error("host build copy failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

int cukd_compute_bounds(void *d_bounds, const void *d_points, int32_t n,
                        const cukd_layout *layout, hipStream_t stream)
{
  g_last_error[0] = 0;
  if (int rc = check_layout(layout)) return rc;
  if (!d_bounds || (!d_points && n > 0)) {
    set_last_error("null pointer");
    return CUKD_ERR_INVALID;
  }
  if (n < 1) {
    set_last_error("computeBounds of an empty set");
    return CUKD_ERR_INVALID;
  }
  return build_bounds_only(d_bounds, d_points, n, *layout, stream);
}

// ---------------------------------------------------------------------------
// query launches: Morton ordering, workgroup size and the autotuner
// ---------------------------------------------------------------------------

namespace {

struct TuneKey {
  int D, trav, list, k, bucket, forest;
  bool operator<(const TuneKey &o) const
  {
    return memcmp(this, &o, sizeof(TuneKey)) < 0;
  }
};
struct TuneVal {
  int block, sort, lanes;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

int log2_bucket(int n)
{
  int b = 0;
  while (b < 30 && (1 << (b + 1)) <= n) b++;
  return b;
}

}  // namespace

// Launches the query in a (order unset).  CUKD_QUERY_SORT: Morton-ordered
// launch.  CUKD_QUERY_AUTOTUNE: the first call per (dims, traversal, list,
// k, log2 nq, forest) times workgroup sizes 64/128/256 with and without
// ordering, and one or two queries per lane where the kernel allows two,
// on these very inputs (every variant writes identical results),
// synchronising the stream once, and caches the fastest; later calls use
// the cache.  repeatable = false (multi-round lists) only reads the cache.
// order_ws (nullable): caller-owned ordering scratch of order_ws_bytes.
static int run_queries(QueryArgs a, int D, int traversal, int flags, const MemResource &mr,
                       void *order_ws, size_t order_ws_bytes, bool repeatable,
                       hipStream_t stream, const char *what)
{
  TuneVal cfg{128, (flags & CUKD_QUERY_SORT) ? 1 : 0, (flags & CUKD_QUERY_ONE_PER_LANE) ? 1 : 2};
  // two queries per lane: stack-based, registers-only lists, round-robin
  // dims, node offsets within 32 bits
  const bool two_ok = traversal == CUKD_STACK_BASED && a.t.dim_offset < 0 &&
                      (a.fq.seg_off ? a.tree_bytes : (uint64_t)a.t.n * (uint64_t)a.t.stride) <=
                          0xffffffffull &&
                      (a.list_kind < 0 || (a.list_kind == CUKD_LIST_FIXED &&
                                           (a.k == 1 || a.k == 4 || a.k == 8 || a.k == 16)));
  bool tune = false;
  TuneKey key;
  memset(&key, 0, sizeof(key));
  key.D = D;
  key.trav = traversal;
  key.list = a.list_kind;
  key.k = a.k;
  key.bucket = log2_bucket(a.nq);
  key.forest = a.fq.seg_off != nullptr;
  if (flags & CUKD_QUERY_AUTOTUNE) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) cfg = it->second;
    else tune = repeatable;
  }
  void *otmp = nullptr;
  bool owned = false;
  if (cfg.sort || tune) {
    const size_t need = query_order_temp_bytes(a.nq);
    if (order_ws && order_ws_bytes >= need) {
      otmp = order_ws;
    } else {
      if (mr.alloc(&otmp, need, stream) != 0) return CUKD_ERR_ALLOC;
      owned = true;
    }
  }
  hipError_t e = hipSuccess;
  if (tune) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float best = 3.4e38f;
    TuneVal bestv = cfg;
    const int blocks[3] = {64, 128, 256};
    for (int lanes = 1; lanes <= (two_ok ? 2 : 1) && e == hipSuccess; lanes++)
    for (int sort = 0; sort < 2 && e == hipSuccess; sort++)
      for (int bi = 0; bi < 3 && e == hipSuccess; bi++) {
        QueryArgs v = a;
        v.block = blocks[bi];
        v.lanes = lanes;
        v.stats = nullptr;  // statistics count the final launch only
        for (int rep = 0; rep < 3 && e == hipSuccess; rep++) {
          if (rep == 1) e = hipEventRecord(e0, stream);
          v.order = sort ? make_query_order(otmp, a.queries, a.nq, D, stream) : nullptr;
          if (e == hipSuccess) e = launch_query(D, traversal, v, stream);
        }
        if (e == hipSuccess) e = hipEventRecord(e1, stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms < best) {
          best = ms;
          bestv = TuneVal{blocks[bi], sort, lanes};
        }
      }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e == hipSuccess) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      g_tune[key] = bestv;
      cfg = bestv;
    }
  }
  if (e == hipSuccess) {
    a.block = cfg.block;
    a.lanes = two_ok ? cfg.lanes : 1;
    a.presorted = (flags & CUKD_QUERY_PRESORTED) != 0;
    a.order = cfg.sort ? make_query_order(otmp, a.queries, a.nq, D, stream) : nullptr;
    e = launch_query(D, traversal, a, stream);
  }
  if (owned) mr.release(otmp, stream);
  if (e != hipSuccess) {
    set_last_error("%s launch failed: %s", what, hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

static int check_query_common(int traversal, const void *d_tree, int32_t n,
                              const cukd_layout *layout, const void *d_world_bounds,
                              const float *d_queries, int32_t nq, int32_t *d_ids)
{
  if (int rc = check_layout(layout)) return rc;
  if (layout->scalar_type != CUKD_SCALAR_F32) {
    set_last_error("queries need float coordinates (the reference defines no "
                   "distance math for int points)");
    return CUKD_ERR_UNSUPPORTED;
  }
  if (traversal < CUKD_STACK_BASED || traversal > CUKD_CCT) {
    set_last_error("unknown traversal %d", traversal);
    return CUKD_ERR_INVALID;
  }
  if (n < 0 || nq < 0) { set_last_error("negative size"); return CUKD_ERR_INVALID; }
  if (n >= (1 << 30)) { set_last_error("tree too large"); return CUKD_ERR_INVALID; }
  if (nq > 0 && (!d_queries || !d_ids || (n > 0 && !d_tree))) {
    set_last_error("null pointer");
    return CUKD_ERR_INVALID;
  }
  if (traversal == CUKD_CCT && !d_world_bounds) {
    set_last_error("cct traversal needs world bounds");
    return CUKD_ERR_INVALID;
  }
  return CUKD_OK;
}

static TreeView make_view(const void *d_tree, int32_t n, const cukd_layout *L)
{
  TreeView t;
  t.base = (const uint8_t *)d_tree;
  t.n = n;
  t.stride = L->stride;
  t.point_offset = L->point_offset;
  t.dim_offset = L->dim_offset;
  t.dim_bytes = L->dim_bytes;
  t.num_dims = L->num_dims;
  return t;
}

int cukd_fcp(int traversal, const void *d_tree, int32_t n, const cukd_layout *layout,
             const void *d_world_bounds, const float *d_queries, int32_t nq, float cutoff,
             float eps, int32_t *d_ids, float *d_dist2, int flags, hipStream_t stream,
             const cukd_mem_resource *mem)
{
  g_last_error[0] = 0;
  if (int rc = check_query_common(traversal, d_tree, n, layout, d_world_bounds, d_queries,
                                  nq, d_ids))
    return rc;
  if (nq == 0) return CUKD_OK;
  MemResource mr{mem};
  QueryArgs a;
  a.t = make_view(d_tree, n, layout);
  a.queries = d_queries;
  a.nq = nq;
  a.world_bounds = (const float *)d_world_bounds;
  a.cutoff = cutoff;
  a.eps = eps;
  a.k = 1;
  a.list_kind = -1;
  a.order = nullptr;
  a.ids = d_ids;
  a.dist2 = d_dist2;
  a.ret = nullptr;
  a.la = ListArgs{nullptr, 0, 0};
  a.fq = ForestQ{nullptr, nullptr, 0, nullptr};
  a.stats = g_stats;
  return run_queries(a, layout->num_dims, traversal, flags, mr, nullptr, 0, true, stream, "fcp");
}

}  // extern "C"

// cukd_knn with the ordering scratch given: order_ws (order_ws_bytes) when
// non-null, else allocated through mr.  tune = false only reads the
// autotuner's cache.
static int knn_entry(int traversal, int list_kind, int k, const void *d_tree, int32_t n,
                     const cukd_layout *layout, const void *d_world_bounds,
                     const float *d_queries, int32_t nq, float cutoff, int32_t *d_ids,
                     float *d_dist2, float *d_max_radius2, int flags, hipStream_t stream,
                     const MemResource &mr, void *order_ws, size_t order_ws_bytes, bool tune)
{
  if (int rc = check_query_common(traversal, d_tree, n, layout, d_world_bounds, d_queries,
                                  nq, d_ids))
    return rc;
  if (k < 1 || k > 64) { set_last_error("k = %d not in 1..64", k); return CUKD_ERR_INVALID; }
  if (list_kind != CUKD_LIST_FIXED && list_kind != CUKD_LIST_HEAP) {
    set_last_error("unknown list kind %d", list_kind);
    return CUKD_ERR_INVALID;
  }
  if (nq > 0 && !d_dist2) { set_last_error("d_dist2 is null"); return CUKD_ERR_INVALID; }
  if (nq == 0) return CUKD_OK;
  QueryArgs a;
  a.t = make_view(d_tree, n, layout);
  a.queries = d_queries;
  a.nq = nq;
  a.world_bounds = (const float *)d_world_bounds;
  a.cutoff = cutoff;
  a.eps = 0.f;
  a.k = k;
  a.list_kind = list_kind;
  a.order = nullptr;
  a.ids = d_ids;
  a.dist2 = d_dist2;
  a.ret = d_max_radius2;
  a.la = ListArgs{nullptr, 0, 0};
  a.fq = ForestQ{nullptr, nullptr, 0, nullptr};
  a.stats = g_stats;
  return run_queries(a, layout->num_dims, traversal, flags, mr, order_ws, order_ws_bytes, tune,
                     stream, "knn");
}

// ---------------------------------------------------------------------------
// allocation-free kNN (cukd_knn_noalloc): the Morton-ordering scratch is a
// file-scope device array, one copy per GPU, held until the process ends
// ---------------------------------------------------------------------------

enum : int32_t { NOALLOC_MAX_QUERIES = 1 << 22 };
static constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)52 << 20;

static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, uint8_t *> g_scratch_ptr;  // device -> its copy of 'scratch'

// The current device's address of 'scratch', looked up once per device.
// (The array's name in host code is only a stand-in for the symbol.)
static int scratch_ptr(uint8_t **p)
{
  int dev = 0;
  CUKD_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    if (query_order_temp_bytes(NOALLOC_MAX_QUERIES) > NOALLOC_SCRATCH_BYTES) {
      set_last_error("internal: the static scratch is smaller than %d queries need",
                     NOALLOC_MAX_QUERIES);
      return CUKD_ERR_INVALID;
    }
    void *a = nullptr;
    const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !a) {
      set_last_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return CUKD_ERR_HIP;
    }
    it = g_scratch_ptr.emplace(dev, (uint8_t *)a).first;
  }
  *p = it->second;
  return CUKD_OK;
}

// a memory resource that refuses every request: no path of
// cukd_knn_noalloc may allocate, and one that tried would fail loudly
static int refuse_malloc(void *, void **ptr, size_t, hipStream_t)
{
  *ptr = nullptr;
  return 1;
}
static int refuse_free(void *, void *, hipStream_t) { return 0; }

extern "C" {

int cukd_knn_noalloc(int traversal, int list_kind, int k, const void *d_tree, int32_t n,
                     const cukd_layout *layout, const void *d_world_bounds,
                     const float *d_queries, int32_t nq, float cutoff, int32_t *d_ids,
                     float *d_dist2, float *d_max_radius2, int flags, hipStream_t stream)
{
  g_last_error[0] = 0;
  if (nq > NOALLOC_MAX_QUERIES) {
    set_last_error("nq = %d exceeds the %d queries that cukd_knn_noalloc's static scratch "
                   "holds",
                   nq, NOALLOC_MAX_QUERIES);
    return CUKD_ERR_INVALID;
  }
  uint8_t *ws = nullptr;
  if (int rc = scratch_ptr(&ws)) return rc;
  static const cukd_mem_resource refuse = {nullptr, refuse_malloc, refuse_free};
  MemResource mr{&refuse};
  return knn_entry(traversal, list_kind, k, d_tree, n, layout, d_world_bounds, d_queries, nq,
                   cutoff, d_ids, d_dist2, d_max_radius2, flags, stream, mr, ws,
                   NOALLOC_SCRATCH_BYTES, false);
}

int cukd_knn(int traversal, int list_kind, int k, const void *d_tree, int32_t n,
             const cukd_layout *layout, const void *d_world_bounds, const float *d_queries,
             int32_t nq, float cutoff, int32_t *d_ids, float *d_dist2, float *d_max_radius2,
             int flags, hipStream_t stream, const cukd_mem_resource *mem)
{
  g_last_error[0] = 0;
  MemResource mr{mem};
  return knn_entry(traversal, list_kind, k, d_tree, n, layout, d_world_bounds, d_queries, nq,
                   cutoff, d_ids, d_dist2, d_max_radius2, flags, stream, mr, nullptr, 0, true);
}

int cukd_knn_accumulate(int traversal, int k, const void *d_tree, int32_t n,
                        const cukd_layout *layout, const void *d_world_bounds,
                        const float *d_queries, int32_t nq, uint64_t *d_lists, int init,
                        float cutoff, int32_t id_offset, int flags, hipStream_t stream,
                        const cukd_mem_resource *mem)
{
  g_last_error[0] = 0;
  if (nq > 0 && !d_lists) { set_last_error("d_lists is null"); return CUKD_ERR_INVALID; }
  if (int rc = check_query_common(traversal, d_tree, n, layout, d_world_bounds, d_queries,
                                  nq, (int32_t *)d_lists))
    return rc;
  if (k < 1 || k > 64) { set_last_error("k = %d not in 1..64", k); return CUKD_ERR_INVALID; }
  if (nq == 0) return CUKD_OK;
  MemResource mr{mem};
  QueryArgs a;
  a.t = make_view(d_tree, n, layout);
  a.queries = d_queries;
  a.nq = nq;
  a.world_bounds = (const float *)d_world_bounds;
  a.cutoff = cutoff;
  a.eps = 0.f;
  a.k = k;
  a.list_kind = LIST_ACCUM;
  a.order = nullptr;
  a.ids = nullptr;
  a.dist2 = nullptr;
  a.ret = nullptr;
  a.la = ListArgs{d_lists, init ? 1 : 0, id_offset};
  a.fq = ForestQ{nullptr, nullptr, 0, nullptr};
  a.stats = g_stats;
  // a round updates its lists in place, so it is never re-run for tuning
  return run_queries(a, layout->num_dims, traversal, flags, mr, nullptr, 0, false, stream,
                     "knn accumulate");
}

int cukd_query_order(const float *d_queries, int32_t nq, int32_t num_dims, int32_t *d_order,
                     hipStream_t stream, const cukd_mem_resource *mem)
{
  g_last_error[0] = 0;
  if (nq < 0 || num_dims < 1 || num_dims > CUKD_MAX_DIMS) {
    set_last_error("bad query count or dims");
    return CUKD_ERR_INVALID;
  }
  if (nq == 0) return CUKD_OK;
  if (!d_queries || !d_order) { set_last_error("null pointer"); return CUKD_ERR_INVALID; }
  MemResource mr{mem};
  void *tmp = nullptr;
  if (mr.alloc(&tmp, query_order_temp_bytes(nq), stream) != 0) return CUKD_ERR_ALLOC;
  const int32_t *order = make_query_order(tmp, d_queries, nq, num_dims, stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_order, order, (size_t)nq * 4, hipMemcpyDeviceToDevice, stream);
  mr.release(tmp, stream);
  if (e != hipSuccess) {
    set_last_error("query order failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

int cukd_knn_extract(const uint64_t *d_lists, int32_t nq, int k, int32_t *d_ids,
                     float *d_dist2, float *d_radius, hipStream_t stream)
{
  g_last_error[0] = 0;
  if (k < 1 || k > 64) { set_last_error("k = %d not in 1..64", k); return CUKD_ERR_INVALID; }
  if (nq < 0 || (nq > 0 && !d_lists)) { set_last_error("bad lists"); return CUKD_ERR_INVALID; }
  if (nq == 0) return CUKD_OK;
  const hipError_t e = launch_knn_extract(d_lists, nq, k, d_ids, d_dist2, d_radius, stream);
  if (e != hipSuccess) {
    set_last_error("knn extract launch failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

int cukd_spatial_build(cukd_spatial_tree *tree, const void *d_points, int32_t n,
                       const cukd_layout *layout, int leaf_threshold, hipStream_t stream,
                       const cukd_mem_resource *mem)
{
  g_last_error[0] = 0;
  if (int rc = check_layout(layout)) return rc;
  if (!tree) { set_last_error("tree is null"); return CUKD_ERR_INVALID; }
  if (n < 0 || (n > 0 && !d_points)) { set_last_error("bad points"); return CUKD_ERR_INVALID; }
  if (leaf_threshold > 65535) {
    set_last_error("leaf threshold %d exceeds the 16-bit leaf count", leaf_threshold);
    return CUKD_ERR_INVALID;
  }
  if (int rc = screen_nan(d_points, n, *layout, stream)) return rc;
  MemResource mr{mem};
  return spatial_build_impl(tree, d_points, n, *layout, leaf_threshold, stream, mr);
}

int cukd_spatial_free(cukd_spatial_tree *tree, hipStream_t stream, const cukd_mem_resource *mem)
{
  g_last_error[0] = 0;
  if (!tree) { set_last_error("tree is null"); return CUKD_ERR_INVALID; }
  MemResource mr{mem};
  mr.release(tree->nodes, stream);
  mr.release(tree->prim_ids, stream);
  tree->nodes = nullptr;
  tree->prim_ids = nullptr;
  tree->num_nodes = 0;
  tree->num_prims = 0;
  return CUKD_OK;
}

static int spatial_query(int traversal, int list_kind, int k, const cukd_spatial_tree *tree,
                         const cukd_layout *layout, const float *d_queries, int32_t nq,
                         float cutoff, int32_t *d_ids, float *d_dist2, float *d_ret,
                         hipStream_t stream)
{
  if (int rc = check_layout(layout)) return rc;
  if (layout->scalar_type != CUKD_SCALAR_F32) {
    set_last_error("spatial queries need float coordinates (the reference defines no "
                   "distance math for int points)");
    return CUKD_ERR_UNSUPPORTED;
  }
  if (traversal != CUKD_STACK_BASED && traversal != CUKD_CCT) {
    set_last_error("spatial trees support stack-based and cct traversal");
    return CUKD_ERR_INVALID;
  }
  if (!tree || nq < 0) { set_last_error("bad arguments"); return CUKD_ERR_INVALID; }
  if (nq == 0) return CUKD_OK;
  if (!tree->nodes || !d_queries || !d_ids) { set_last_error("null pointer"); return CUKD_ERR_INVALID; }
  SpatialArgs a;
  a.nodes = tree->nodes;
  a.prim_ids = tree->prim_ids;
  a.data = make_view(tree->data, tree->num_prims, layout);
  for (int i = 0; i < 2 * CUKD_MAX_DIMS; i++) a.bounds.v[i] = tree->bounds[i];
  a.queries = d_queries;
  a.nq = nq;
  a.cutoff = cutoff;
  a.k = k;
  a.list_kind = list_kind;
  a.ids = d_ids;
  a.dist2 = d_dist2;
  a.ret = d_ret;
  a.stats = g_stats;
  const hipError_t e = launch_spatial_query(layout->num_dims, traversal, a, stream);
  if (e != hipSuccess) {
    set_last_error("spatial query launch failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

int cukd_spatial_fcp(int traversal, const cukd_spatial_tree *tree, const cukd_layout *layout,
                     const float *d_queries, int32_t nq, float cutoff, int32_t *d_ids,
                     float *d_dist2, hipStream_t stream)
{
  g_last_error[0] = 0;
  return spatial_query(traversal, -1, 1, tree, layout, d_queries, nq, cutoff, d_ids, d_dist2,
                       nullptr, stream);
}

int cukd_spatial_knn(int traversal, int list_kind, int k, const cukd_spatial_tree *tree,
                     const cukd_layout *layout, const float *d_queries, int32_t nq,
                     float cutoff, int32_t *d_ids, float *d_dist2, float *d_max_radius2,
                     hipStream_t stream)
{
  g_last_error[0] = 0;
  if (k < 1 || k > 64) { set_last_error("k = %d not in 1..64", k); return CUKD_ERR_INVALID; }
  if (list_kind != CUKD_LIST_FIXED && list_kind != CUKD_LIST_HEAP) {
    set_last_error("unknown list kind %d", list_kind);
    return CUKD_ERR_INVALID;
  }
  if (nq > 0 && !d_dist2) { set_last_error("d_dist2 is null"); return CUKD_ERR_INVALID; }
  return spatial_query(traversal, list_kind, k, tree, layout, d_queries, nq, cutoff, d_ids,
                       d_dist2, d_max_radius2, stream);
}

// ---------------------------------------------------------------------------
// forests: batched build and queries over many trees
// ---------------------------------------------------------------------------

int cukd_forest_create(cukd_forest **forest, const int64_t *h_offsets, int32_t num_trees,
                       const cukd_layout *layout, hipStream_t stream,
                       const cukd_mem_resource *mem)
{
  g_last_error[0] = 0;
  if (!forest) { set_last_error("forest is null"); return CUKD_ERR_INVALID; }
  *forest = nullptr;
  if (int rc = check_layout(layout)) return rc;
  if (num_trees < 1 || !h_offsets) {
    set_last_error("need num_trees >= 1 and an offsets array");
    return CUKD_ERR_INVALID;
  }
  if (h_offsets[0] != 0) { set_last_error("offsets must start at 0"); return CUKD_ERR_INVALID; }
  for (int t = 0; t < num_trees; t++) {
    const int64_t n = h_offsets[t + 1] - h_offsets[t];
    if (n < 0 || n >= (1 << 30)) {
      set_last_error("tree %d has %lld points (must be 0 .. 2^30-1)", t, (long long)n);
      return CUKD_ERR_INVALID;
    }
  }
  if (h_offsets[num_trees] >= ((int64_t)1 << 31)) {
    set_last_error("forest of %lld points exceeds 2^31-1", (long long)h_offsets[num_trees]);
    return CUKD_ERR_INVALID;
  }
  cukd_forest *f = new cukd_forest();
  f->has_mem = mem != nullptr;
  if (mem) f->mem = *mem;
  std::vector<long long> off(num_trees + 1);
  for (int t = 0; t <= num_trees; t++) off[t] = (long long)h_offsets[t];
  make_forest_plan(f->plan, off.data(), num_trees, *layout, true);
  const size_t off_trees = (f->plan.total + 255) & ~(size_t)255;
  f->bytes = off_trees + (size_t)(num_trees + 1) * 4;
  MemResource mr{f->has_mem ? &f->mem : nullptr};
  if (mr.alloc((void **)&f->ws, f->bytes, stream) != 0) {
    delete f;
    return CUKD_ERR_ALLOC;
  }
  f->tree_off = (int32_t *)(f->ws + off_trees);
  std::vector<uint8_t> img;
  size_t begin = 0;
  forest_tables_image(f->plan, img, begin);
  std::vector<int32_t> toff(num_trees + 1);
  for (int t = 0; t <= num_trees; t++) toff[t] = (int32_t)h_offsets[t];
  hipError_t e = hipSuccess;
  if (!img.empty()) e = hipMemcpyAsync(f->ws + begin, img.data(), img.size(), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(f->tree_off, toff.data(), toff.size() * 4, hipMemcpyHostToDevice, stream);
  // the host images go out of scope here
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    set_last_error("forest set-up failed: %s", hipGetErrorString(e));
    mr.release(f->ws, stream);
    delete f;
    return CUKD_ERR_HIP;
  }
  *forest = f;
  return CUKD_OK;
}

int cukd_forest_workspace_bytes(const cukd_forest *forest, size_t *bytes)
{
  g_last_error[0] = 0;
  if (!forest || !bytes) { set_last_error("null argument"); return CUKD_ERR_INVALID; }
  *bytes = forest->bytes;
  return CUKD_OK;
}

int cukd_forest_build(cukd_forest *forest, void *d_points, void *d_world_bounds,
                      hipStream_t stream)
{
  g_last_error[0] = 0;
  if (!forest) { set_last_error("forest is null"); return CUKD_ERR_INVALID; }
  if (forest->plan.ntotal > 0 && !d_points) {
    set_last_error("d_points is null");
    return CUKD_ERR_INVALID;
  }
  if (forest->plan.expl && !d_world_bounds) {
    set_last_error("asked to build k-d trees over nodes with explicit dims, "
                   "but no memory for world bounds provided");
    return CUKD_ERR_NO_BOUNDS;
  }
  return forest_build_run(forest->plan, d_points, d_world_bounds, forest->ws, stream);
}

int cukd_forest_reserve_queries(cukd_forest *forest, int32_t max_queries, hipStream_t stream)
{
  g_last_error[0] = 0;
  if (!forest || max_queries < 0) { set_last_error("bad argument"); return CUKD_ERR_INVALID; }
  const size_t need = query_order_temp_bytes(max_queries);
  std::lock_guard<std::mutex> lk(forest->order_mu);
  if (need <= forest->order_ws_bytes && stream == forest->order_stream) return CUKD_OK;
  const size_t bytes = need > forest->order_ws_bytes ? need : forest->order_ws_bytes;
  MemResource mr{forest->has_mem ? &forest->mem : nullptr};
  // the old scratch was only ever used on the old stream: free it there,
  // after those queries
  if (forest->order_ws) mr.release(forest->order_ws, forest->order_stream);
  forest->order_ws = nullptr;
  forest->order_ws_bytes = 0;
  forest->order_stream = stream;
  if (mr.alloc(&forest->order_ws, bytes, stream) != 0) return CUKD_ERR_ALLOC;
  forest->order_ws_bytes = bytes;
  return CUKD_OK;
}

int cukd_forest_destroy(cukd_forest *forest, hipStream_t stream)
{
  g_last_error[0] = 0;
  if (!forest) return CUKD_OK;
  MemResource mr{forest->has_mem ? &forest->mem : nullptr};
  if (forest->order_ws) {
    // queries on the reserving stream may still be using it
    if (forest->order_stream != stream) (void)hipStreamSynchronize(forest->order_stream);
    mr.release(forest->order_ws, stream);
  }
  mr.release(forest->ws, stream);
  delete forest;
  return CUKD_OK;
}

static int forest_query(int traversal, int list_kind, int k, const cukd_forest *forest,
                        const void *d_points, const void *d_world_bounds,
                        const float *d_queries, int32_t nq, int32_t num_segments,
                        const int32_t *d_seg_offsets, const int32_t *d_seg_tree, float cutoff,
                        float eps, int32_t *d_ids, float *d_dist2, float *d_ret, int flags,
                        hipStream_t stream)
{
  if (!forest) { set_last_error("forest is null"); return CUKD_ERR_INVALID; }
  const cukd_layout &L = forest->plan.L;
  if (int rc = check_query_common(traversal, d_points, (int32_t)(forest->plan.ntotal > 0), &L,
                                  d_world_bounds, d_queries, nq, d_ids))
    return rc;
  if (nq == 0) return CUKD_OK;
  if (num_segments < 1 || !d_seg_offsets || !d_seg_tree) {
    set_last_error("need num_segments >= 1 with segment offsets and trees");
    return CUKD_ERR_INVALID;
  }
  QueryArgs a;
  a.t = make_view(d_points, 0, &L);
  a.queries = d_queries;
  a.nq = nq;
  a.world_bounds = (const float *)d_world_bounds;
  a.cutoff = cutoff;
  a.eps = eps;
  a.k = k;
  a.list_kind = list_kind;
  a.order = nullptr;
  a.ids = d_ids;
  a.dist2 = d_dist2;
  a.ret = d_ret;
  a.la = ListArgs{nullptr, 0, 0};
  a.fq = ForestQ{d_seg_offsets, d_seg_tree, num_segments, forest->tree_off};
  a.tree_bytes = (uint64_t)forest->plan.ntotal * (uint64_t)L.stride;
  a.max_n = (int)forest->plan.max_n;
  a.stats = g_stats;
  MemResource mr{forest->has_mem ? &forest->mem : nullptr};
  // the reserved ordering scratch serves queries on its own stream only;
  // sorted queries on any other stream get scratch of their own, so two
  // streams never share one Morton order
  cukd_forest *f = const_cast<cukd_forest *>(forest);
  std::unique_lock<std::mutex> lk(f->order_mu);
  const bool own = f->order_ws && stream == f->order_stream;
  void *ws = own ? f->order_ws : nullptr;
  const size_t ws_bytes = own ? f->order_ws_bytes : 0;
  if (!own) lk.unlock();
  return run_queries(a, L.num_dims, traversal, flags, mr, ws, ws_bytes, true, stream,
                     "forest query");
}

int cukd_forest_fcp(int traversal, const cukd_forest *forest, const void *d_points,
                    const void *d_world_bounds, const float *d_queries, int32_t nq,
                    int32_t num_segments, const int32_t *d_seg_offsets,
                    const int32_t *d_seg_tree, float cutoff, float eps, int32_t *d_ids,
                    float *d_dist2, int flags, hipStream_t stream)
{
  g_last_error[0] = 0;
  return forest_query(traversal, -1, 1, forest, d_points, d_world_bounds, d_queries, nq,
                      num_segments, d_seg_offsets, d_seg_tree, cutoff, eps, d_ids, d_dist2,
                      nullptr, flags, stream);
}

int cukd_forest_knn(int traversal, int list_kind, int k, const cukd_forest *forest,
                    const void *d_points, const void *d_world_bounds, const float *d_queries,
                    int32_t nq, int32_t num_segments, const int32_t *d_seg_offsets,
                    const int32_t *d_seg_tree, float cutoff, int32_t *d_ids, float *d_dist2,
                    float *d_max_radius2, int flags, hipStream_t stream)
{
  g_last_error[0] = 0;
  if (k < 1 || k > 64) { set_last_error("k = %d not in 1..64", k); return CUKD_ERR_INVALID; }
  if (list_kind != CUKD_LIST_FIXED && list_kind != CUKD_LIST_HEAP) {
    set_last_error("unknown list kind %d", list_kind);
    return CUKD_ERR_INVALID;
  }
  if (nq > 0 && !d_dist2) { set_last_error("d_dist2 is null"); return CUKD_ERR_INVALID; }
  return forest_query(traversal, list_kind, k, forest, d_points, d_world_bounds, d_queries, nq,
                      num_segments, d_seg_offsets, d_seg_tree, cutoff, 0.f, d_ids, d_dist2,
                      d_max_radius2, flags, stream);
}

}  // extern "C"
