/* cukd_oracle.c -- CPU restatement of cudaKDTree's build-and-query path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity oracle and CPU baseline.  Only the
 * test suite, the smoke check and bench.py's cpu_baseline leg load it.
 *
 * Floating point: compiled with -ffp-contract=off, so every distance is
 * evaluated as the reference host path does, d = a - b per component and
 * dist2 = ((d0*d0 + d1*d1) + d2*d2) + ... left to right, with no fused
 * multiply-add (cukd-math.h:100-112, 210-213).
 *
 * Citations are file:line in the ingowald/cudaKDTree v1.0.1 tree.
 */
#include "cukd_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OC_MAX_DIMS 8

/* ------------------------------------------------------------------ */
/* element access (data.h:56-116, cukd-math.h:298-331)                 */
/* ------------------------------------------------------------------ */

static inline float get_f(const uint8_t *e, const oc_layout *L, int d)
{
  float f;
  memcpy(&f, e + L->point_offset + 4 * d, 4);
  return f;
}

static inline int32_t get_i(const uint8_t *e, const oc_layout *L, int d)
{
  int32_t i;
  memcpy(&i, e + L->point_offset + 4 * d, 4);
  return i;
}

static inline int get_dim(const uint8_t *e, const oc_layout *L)
{
  const uint8_t *p = e + L->dim_offset;
  if (L->dim_bytes == 1) return (int)p[0];
  if (L->dim_bytes == 2) { uint16_t v; memcpy(&v, p, 2); return (int)v; }
  int32_t v; memcpy(&v, p, 4); return (int)v;
}

static inline void set_dim(uint8_t *e, const oc_layout *L, int dim)
{
  uint8_t *p = e + L->dim_offset;
  if (L->dim_bytes == 1) { p[0] = (uint8_t)dim; return; }
  if (L->dim_bytes == 2) { uint16_t v = (uint16_t)dim; memcpy(p, &v, 2); return; }
  int32_t v = dim; memcpy(p, &v, 4);
}

/* coord(a)[d] < coord(b)[d] in the scalar type (ZipCompare's '<',
 * builder_thrust.h:401-404). */
static inline int coord_lt(const uint8_t *a, const uint8_t *b,
                           const oc_layout *L, int d)
{
  if (L->scalar_type == OC_F32) return get_f(a, L, d) < get_f(b, L, d);
  return get_i(a, L, d) < get_i(b, L, d);
}

/* box layout: box_t<point_t> = {point_t lower, upper} (box.h:34-82),
 * i.e. D scalars of lower followed by D scalars of upper. */
static inline float boxf(const void *box, int which, int d, int D)
{ return ((const float *)box)[which * D + d]; }
static inline int32_t boxi(const void *box, int which, int d, int D)
{ return ((const int32_t *)box)[which * D + d]; }

/* ------------------------------------------------------------------ */
/* tree index math (helpers.h:76-258)                                   */
/* ------------------------------------------------------------------ */

/* BinaryTree::levelOf (helpers.h:84-97) */
int oc_level_of(int64_t node)
{
  return 63 - __builtin_clzll((unsigned long long)(node + 1));
}

/* BinaryTree::numLevelsFor (helpers.h:98-101) */
int oc_num_levels_for(int64_t n) { return oc_level_of(n - 1) + 1; }

/* ArbitraryBinaryTree::numNodesInSubtree (helpers.h:162-178): nodes of a
 * full subtree under 'node', minus the ones whose index is >= n on the
 * deepest level. */
int64_t oc_subtree_size(int64_t node, int64_t n)
{
  if (node >= n) return 0;
  const int levels_tree = oc_num_levels_for(n);
  const int levels_sub = levels_tree - oc_level_of(node);
  const int64_t on_last = (int64_t)1 << (levels_sub - 1);
  const int64_t full = ((int64_t)1 << levels_sub) - 1;
  /* one past the last node index of the subtree's deepest level */
  const int64_t end_last = (node + 2) * on_last - 1;
  int64_t missing = end_last - n;
  if (missing < 0) missing = 0;
  if (missing > on_last) missing = on_last;
  return full - missing;
}

/* ArrayLayoutInStep::segmentBegin (helpers.h:206-233) */
int64_t oc_segment_begin(int step, int64_t n, int64_t subtree)
{
  const int64_t settled = ((int64_t)1 << step) - 1;
  const int levels_total = oc_num_levels_for(n);
  const int remaining = levels_total - step;
  const int64_t earlier = subtree - settled;
  const int64_t left_if_full = earlier * (((int64_t)1 << remaining) - 1);
  const int64_t left_on_last_if_full = earlier * ((int64_t)1 << (remaining - 1));
  const int64_t total_on_last = n - (((int64_t)1 << (levels_total - 1)) - 1);
  const int64_t really = total_on_last < left_on_last_if_full
                             ? total_on_last : left_on_last_if_full;
  return settled + left_if_full - (left_on_last_if_full - really);
}

/* ArrayLayoutInStep::pivotPosOf / sizeOfLeftSubtreeOf (helpers.h:235-249) */
int64_t oc_pivot_pos(int step, int64_t n, int64_t subtree)
{
  const int64_t left = 2 * subtree + 1;
  const int64_t left_size = left >= n ? 0 : oc_subtree_size(left, n);
  return oc_segment_begin(step, n, subtree) + left_size;
}

/* ------------------------------------------------------------------ */
/* bounds (builder_common.h:166-174, box.h:54-68, 110-125)              */
/* ------------------------------------------------------------------ */

void oc_compute_bounds(const void *points, int64_t n, const oc_layout *L,
                       void *box_out)
{
  const int D = L->num_dims;
  const uint8_t *base = (const uint8_t *)points;
  if (L->scalar_type == OC_F32) {
    float *b = (float *)box_out;
    for (int d = 0; d < D; d++) { b[d] = INFINITY; b[D + d] = -INFINITY; }
    for (int64_t i = 0; i < n; i++) {
      const uint8_t *e = base + i * L->stride;
      for (int d = 0; d < D; d++) {
        const float f = get_f(e, L, d);
        /* lower = min(lower,p), upper = max(upper,p) with std::min/max */
        if (f < b[d]) b[d] = f;
        if (b[D + d] < f) b[D + d] = f;
      }
    }
  } else {
    int32_t *b = (int32_t *)box_out;
    for (int d = 0; d < D; d++) { b[d] = INT_MAX; b[D + d] = INT_MIN; }
    for (int64_t i = 0; i < n; i++) {
      const uint8_t *e = base + i * L->stride;
      for (int d = 0; d < D; d++) {
        const int32_t v = get_i(e, L, d);
        if (v < b[d]) b[d] = v;
        if (b[D + d] < v) b[D + d] = v;
      }
    }
  }
}

/* box_t::widestDimension (box.h:110-125): ties pick the later dim. */
int oc_widest_dim(const void *box, const oc_layout *L)
{
  const int D = L->num_dims;
  int d_best = 0;
  if (L->scalar_type == OC_F32) {
    float w_best = 0.f;
    for (int d = 0; d < D; d++) {
      const float w = boxf(box, 1, d, D) - boxf(box, 0, d, D);
      if (w < w_best) continue;
      w_best = w;
      d_best = d;
    }
  } else {
    int64_t w_best = 0;
    for (int d = 0; d < D; d++) {
      const int64_t w = (int64_t)boxi(box, 1, d, D) - (int64_t)boxi(box, 0, d, D);
      if (w < w_best) continue;
      w_best = w;
      d_best = d;
    }
  }
  return d_best;
}

/* ------------------------------------------------------------------ */
/* host builder (builder_thrust.h:514-601)                              */
/* ------------------------------------------------------------------ */

typedef struct { uint32_t tag, elem; } rec_t;

typedef struct {
  const oc_layout *L;
  const uint8_t *elems;
  int dim; /* round-robin dim of this step; ignored for explicit dims */
} zip_ctx;

/* ZipCompare::operator() (builder_thrust.h:380-407): (tag, coord[dim])
 * with dim taken from the first operand for explicit-dim data. */
static inline int zip_less(const zip_ctx *c, rec_t a, rec_t b)
{
  if (a.tag != b.tag) return a.tag < b.tag;
  const uint8_t *ea = c->elems + (size_t)a.elem * c->L->stride;
  const uint8_t *eb = c->elems + (size_t)b.elem * c->L->stride;
  const int dim = c->L->dim_offset >= 0 ? get_dim(ea, c->L) : c->dim;
  return coord_lt(ea, eb, c->L, dim);
}

/* Stable merge sort.  thrust::sort with a user comparator is a stable
 * merge sort in both the host and device backends, so any stable sort
 * with the same comparator yields the same permutation. */
static void stable_sort(rec_t *a, rec_t *tmp, int64_t n, const zip_ctx *c)
{
  const int64_t RUN = 32;
  for (int64_t lo = 0; lo < n; lo += RUN) {
    const int64_t hi = lo + RUN < n ? lo + RUN : n;
    for (int64_t i = lo + 1; i < hi; i++) {
      rec_t v = a[i];
      int64_t j = i;
      while (j > lo && zip_less(c, v, a[j - 1])) { a[j] = a[j - 1]; j--; }
      a[j] = v;
    }
  }
  rec_t *src = a, *dst = tmp;
  for (int64_t width = RUN; width < n; width *= 2) {
    for (int64_t lo = 0; lo < n; lo += 2 * width) {
      const int64_t mid = lo + width < n ? lo + width : n;
      const int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
      int64_t i = lo, j = mid, k = lo;
      while (i < mid && j < hi) {
        if (zip_less(c, src[j], src[i])) dst[k++] = src[j++];
        else dst[k++] = src[i++];
      }
      while (i < mid) dst[k++] = src[i++];
      while (j < hi) dst[k++] = src[j++];
    }
    rec_t *t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, (size_t)n * sizeof(rec_t));
}

/* findBounds (builder_common.h:181-219): walk up from 'subtree', clipping
 * the world box at every ancestor's split plane. */
static void find_bounds(int64_t subtree, const void *world, const rec_t *recs,
                        const uint8_t *elems, const oc_layout *L,
                        void *box_out)
{
  const int D = L->num_dims;
  memcpy(box_out, world, (size_t)2 * D * 4);
  int64_t curr = subtree;
  while (curr > 0) {
    const int64_t parent = (curr + 1) / 2 - 1;
    const uint8_t *pn = elems + (size_t)recs[parent].elem * L->stride;
    const int pdim = get_dim(pn, L);
    if (L->scalar_type == OC_F32) {
      float *b = (float *)box_out;
      const float s = get_f(pn, L, pdim);
      if (curr & 1) { /* left child: upper = min(split, upper) */
        const float u = b[D + pdim];
        b[D + pdim] = (u < s) ? u : s;
      } else {        /* right child: lower = max(split, lower) */
        const float l = b[pdim];
        b[pdim] = (s < l) ? l : s;
      }
    } else {
      int32_t *b = (int32_t *)box_out;
      const int32_t s = get_i(pn, L, pdim);
      if (curr & 1) { const int32_t u = b[D + pdim]; b[D + pdim] = (u < s) ? u : s; }
      else          { const int32_t l = b[pdim];     b[pdim]     = (s < l) ? l : s; }
    }
    curr = parent;
  }
}

int oc_build_host(void *points, int64_t n, const oc_layout *L,
                  void *world_bounds)
{
  if (n < 1) return 0; /* builder_thrust.h:539 */
  const int D = L->num_dims;
  const int explicit_dim = L->dim_offset >= 0;
  if (world_bounds) oc_compute_bounds(points, n, L, world_bounds);
  if (explicit_dim && !world_bounds) return -1;

  const size_t stride = (size_t)L->stride;
  uint8_t *elems = (uint8_t *)malloc((size_t)n * stride);
  rec_t *recs = (rec_t *)malloc((size_t)n * sizeof(rec_t));
  rec_t *tmp = (rec_t *)malloc((size_t)n * sizeof(rec_t));
  if (!elems || !recs || !tmp) { free(elems); free(recs); free(tmp); return -2; }
  memcpy(elems, points, (size_t)n * stride);
  for (int64_t i = 0; i < n; i++) { recs[i].tag = 0; recs[i].elem = (uint32_t)i; }

  /* host_chooseInitialDim (builder_thrust.h:158-168) */
  if (explicit_dim) {
    const int d0 = oc_widest_dim(world_bounds, L);
    for (int64_t i = 0; i < n; i++) set_dim(elems + i * stride, L, d0);
  }

  const int num_levels = oc_num_levels_for(n);
  const int deepest = num_levels - 1;
  uint8_t boxbuf[2 * OC_MAX_DIMS * 4];
  zip_ctx ctx = {L, elems, 0};
  for (int level = 0; level < deepest; level++) {
    ctx.dim = level % D;
    stable_sort(recs, tmp, n, &ctx);
    const int64_t settled = ((int64_t)1 << level) - 1;
    if (explicit_dim) {
      /* host_updateTagsAndSetDims -> updateTagAndSetDim
       * (builder_thrust.h:259-314, 352-374) */
      for (int64_t gid = settled; gid < n; gid++) {
        int64_t subtree = recs[gid].tag;
        find_bounds(subtree, world_bounds, recs, elems, L, boxbuf);
        const int64_t pivot = oc_pivot_pos(level, n, subtree);
        const uint8_t *pe = elems + (size_t)recs[pivot].elem * stride;
        const int pdim = get_dim(pe, L);
        if (gid < pivot) {
          subtree = 2 * subtree + 1;
          memcpy(boxbuf + (size_t)(D + pdim) * 4, pe + L->point_offset + 4 * pdim, 4);
        } else if (gid > pivot) {
          subtree = 2 * subtree + 2;
          memcpy(boxbuf + (size_t)pdim * 4, pe + L->point_offset + 4 * pdim, 4);
        }
        if (gid != pivot)
          set_dim(elems + (size_t)recs[gid].elem * stride, L, oc_widest_dim(boxbuf, L));
        recs[gid].tag = (uint32_t)subtree;
      }
    } else {
      /* host_updateTags -> updateTag (builder_thrust.h:175-211, 242-251) */
      for (int64_t gid = settled; gid < n; gid++) {
        const int64_t subtree = recs[gid].tag;
        const int64_t pivot = oc_pivot_pos(level, n, subtree);
        if (gid < pivot) recs[gid].tag = (uint32_t)(2 * subtree + 1);
        else if (gid > pivot) recs[gid].tag = (uint32_t)(2 * subtree + 2);
      }
    }
  }
  /* final sort: every tag is now a unique node id (builder_thrust.h:597-600) */
  ctx.dim = deepest % D;
  stable_sort(recs, tmp, n, &ctx);

  uint8_t *out = (uint8_t *)points;
  for (int64_t i = 0; i < n; i++)
    memcpy(out + i * stride, elems + (size_t)recs[i].elem * stride, stride);
  free(elems); free(recs); free(tmp);
  return 0;
}

/* ------------------------------------------------------------------ */
/* result types (fcp.h:179-207, knn.h:343-548)                          */
/* ------------------------------------------------------------------ */

typedef struct {
  int kind;          /* 0 = FCP, 1 = fixed list, 2 = heap list */
  int k;
  int id_offset;     /* added to every candidate id (multi-round lists) */
  uint64_t *entry;   /* k entries for kNN */
  int closest_id;    /* FCP */
  float closest_d2;  /* FCP */
} result_t;

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* BaseCandidateList::encode (knn.h:343-345) */
static inline uint64_t enc(float d2, int id)
{ return ((uint64_t)f2u(d2) << 32) | (uint32_t)id; }

static void result_init(result_t *r, float cutoff)
{
  const float r2 = cutoff * cutoff; /* sqr(cutOffRadius) */
  if (r->kind == 0) { r->closest_d2 = r2; r->closest_id = -1; return; }
  for (int i = 0; i < r->k; i++) r->entry[i] = enc(r2, -1);
}

static inline float result_max_r2(const result_t *r)
{
  if (r->kind == 0) return r->closest_d2;
  if (r->kind == 1) return u2f((uint32_t)(r->entry[r->k - 1] >> 32));
  return u2f((uint32_t)(r->entry[0] >> 32));
}

/* FCPResult::processCandidate (fcp.h:191-199), FixedCandidateList::push
 * (knn.h:523-532), HeapCandidateList::push (knn.h:456-483). */
static inline float result_process(result_t *r, int id, float d2)
{
  if (r->kind == 0) {
    if (d2 < r->closest_d2) { r->closest_d2 = d2; r->closest_id = id; }
    return r->closest_d2;
  }
  uint64_t v = enc(d2, id + r->id_offset);
  uint64_t *e = r->entry;
  const int k = r->k;
  if (r->kind == 1) {
    for (int i = 0; i < k; i++) {
      const uint64_t lo = e[i] < v ? e[i] : v;
      const uint64_t hi = e[i] < v ? v : e[i];
      e[i] = lo;
      v = hi;
    }
    return u2f((uint32_t)(e[k - 1] >> 32));
  }
  if (v >= e[0]) return u2f((uint32_t)(e[0] >> 32));
  int pos = 0;
  for (;;) {
    uint64_t largest_val = UINT64_MAX;
    const int c0 = 2 * pos + 1;
    int largest = k;
    if (c0 < k) { largest = c0; largest_val = e[c0]; }
    const int c1 = c0 + 1;
    if (c1 < k && e[c1] > largest_val) { largest = c1; largest_val = e[c1]; }
    if (largest == k || largest_val < v) { e[pos] = v; break; }
    e[pos] = largest_val;
    pos = largest;
  }
  return u2f((uint32_t)(e[0] >> 32));
}

/* ------------------------------------------------------------------ */
/* distance (cukd-math.h:100-112, 210-213)                              */
/* ------------------------------------------------------------------ */

static inline float sqr_dist(const float *a, const float *b, int D)
{
  float d = a[0] - b[0];
  float s = d * d;
  for (int i = 1; i < D; i++) { d = a[i] - b[i]; s = s + d * d; }
  return s;
}

static inline void load_point(const uint8_t *e, const oc_layout *L, float *p)
{
  memcpy(p, e + L->point_offset, (size_t)4 * L->num_dims);
}

/* ------------------------------------------------------------------ */
/* traversals over the left-balanced tree                               */
/* ------------------------------------------------------------------ */

/* traverse_default (traverse-default-stack-based.h:24-108) */
static int64_t traverse_stack_based(result_t *res, const float *q,
                                    const uint8_t *nodes, int64_t n,
                                    const oc_layout *L)
{
  const int D = L->num_dims;
  const int explicit_dim = L->dim_offset >= 0;
  float cull = result_max_r2(res);
  int64_t visits = 0;
  int stack_node[30];
  float stack_d2[30];
  int sp = 0;
  int64_t curr = 0;
  float p[OC_MAX_DIMS];
  for (;;) {
    while (curr < n) {
      const uint8_t *e = nodes + (size_t)curr * L->stride;
      const int dim = explicit_dim ? get_dim(e, L) : oc_level_of(curr) % D;
      visits++;
      load_point(e, L, p);
      const float d2 = sqr_dist(p, q, D);
      cull = result_process(res, (int)curr, d2);
      const float node_c = p[dim];
      const float query_c = q[dim];
      const int left_close = query_c < node_c;
      const int64_t lchild = 2 * curr + 1;
      const int64_t rchild = lchild + 1;
      const int64_t close_c = left_close ? lchild : rchild;
      const int64_t far_c = left_close ? rchild : lchild;
      const float diff = query_c - node_c;
      const float plane2 = diff * diff;
      if (plane2 < cull && far_c < n) {
        stack_node[sp] = (int)far_c;
        stack_d2[sp] = plane2;
        sp++;
      }
      curr = close_c;
    }
    for (;;) {
      if (sp == 0) return visits;
      --sp;
      if (stack_d2[sp] >= cull) continue;
      curr = stack_node[sp];
      break;
    }
  }
}

/* traverse_stack_free (traverse-stack-free.h:24-117) */
static int64_t traverse_stack_free(result_t *res, const float *q,
                                   const uint8_t *nodes, int64_t n,
                                   const oc_layout *L, float eps)
{
  const int D = L->num_dims;
  const int explicit_dim = L->dim_offset >= 0;
  const float eps_err = 1.f + eps;
  float cull = result_max_r2(res);
  int64_t visits = 0;
  int64_t prev = -1, curr = 0;
  float p[OC_MAX_DIMS];
  if (n <= 0) return 0; /* the reference walks off the array for N = 0 */
  for (;;) {
    const int64_t parent = (curr + 1) / 2 - 1;
    if (curr >= n) {
      prev = curr;
      curr = parent;
      continue;
    }
    visits++;
    const uint8_t *e = nodes + (size_t)curr * L->stride;
    load_point(e, L, p);
    const int64_t child = 2 * curr + 1;
    const int from_child = prev >= child;
    if (!from_child) {
      const float d2 = sqr_dist(q, p, D);
      cull = result_process(res, (int)curr, d2);
    }
    const int dim = explicit_dim ? get_dim(e, L) : oc_level_of(curr) % D;
    const float dd = q[dim] - p[dim];
    const int side = dd > 0.f;
    const int64_t close_c = 2 * curr + 1 + side;
    const int64_t far_c = 2 * curr + 2 - side;
    int64_t next;
    if (prev == close_c)
      next = (far_c < n && (dd * dd) * eps_err < cull) ? far_c : parent;
    else if (prev == far_c)
      next = parent;
    else
      next = child < n ? close_c : parent;
    if (next == -1) return visits;
    prev = curr;
    curr = next;
  }
}

/* traverse_cct (traverse-cct.h:25-99) */
static int64_t traverse_cct(result_t *res, const float *q, const float *wb,
                            const uint8_t *nodes, int64_t n,
                            const oc_layout *L)
{
  const int D = L->num_dims;
  const int explicit_dim = L->dim_offset >= 0;
  float cull = result_max_r2(res);
  int64_t visits = 0;
  int stack_node[30];
  float stack_corner[30][OC_MAX_DIMS];
  int sp = 0;
  int64_t node_id = 0;
  float closest[OC_MAX_DIMS] = {0}, p[OC_MAX_DIMS], far_corner[OC_MAX_DIMS];
  /* project(bounds, q) = min(max(q, lower), upper) (box.h:96-102) */
  for (int d = 0; d < D; d++) {
    float v = q[d];
    const float lo = wb[d], hi = wb[D + d];
    v = (v < lo) ? lo : v;
    v = (hi < v) ? hi : v;
    closest[d] = v;
  }
  if (sqr_dist(q, closest, D) > cull) return visits;
  for (;;) {
    if (node_id >= n) {
      for (;;) {
        if (sp == 0) return visits;
        --sp;
        memcpy(closest, stack_corner[sp], sizeof(float) * D);
        if (sqr_dist(closest, q, D) >= cull) continue;
        node_id = stack_node[sp];
        break;
      }
    }
    const uint8_t *e = nodes + (size_t)node_id * L->stride;
    visits++;
    load_point(e, L, p);
    cull = result_process(res, (int)node_id, sqr_dist(p, q, D));
    const int dim = explicit_dim ? get_dim(e, L) : oc_level_of(node_id) % D;
    const float node_c = p[dim];
    const float query_c = q[dim];
    const int left_close = query_c < node_c;
    const int64_t lchild = 2 * node_id + 1;
    const int64_t rchild = lchild + 1;
    memcpy(far_corner, closest, sizeof(float) * D);
    const int64_t far_c = left_close ? rchild : lchild;
    far_corner[dim] = node_c;
    if (far_c < n && sqr_dist(far_corner, q, D) < cull) {
      memcpy(stack_corner[sp], far_corner, sizeof(float) * D);
      stack_node[sp] = (int)far_c;
      sp++;
    }
    node_id = left_close ? lchild : rchild;
  }
}

static int64_t run_traversal(int method, result_t *res, const float *q,
                             const void *wb, const uint8_t *nodes, int64_t n,
                             const oc_layout *L, float eps)
{
  if (method == OC_STACK_BASED) return traverse_stack_based(res, q, nodes, n, L);
  if (method == OC_STACK_FREE) return traverse_stack_free(res, q, nodes, n, L, eps);
  return traverse_cct(res, q, (const float *)wb, nodes, n, L);
}

/* stackBased/stackFree/cct::fcp (fcp.h:210-254) */
void oc_fcp(int method, const void *tree, int64_t n, const oc_layout *L,
            const void *world_bounds, const float *queries, int64_t nq,
            float cutoff, float eps, int32_t *ids, float *dist2,
            int64_t *visits, int num_threads)
{
  const int D = L->num_dims;
#pragma omp parallel for schedule(dynamic, 256) num_threads(num_threads > 0 ? num_threads : 1)
  for (int64_t i = 0; i < nq; i++) {
    result_t r;
    r.kind = 0; r.k = 1; r.entry = NULL; r.id_offset = 0;
    result_init(&r, cutoff);
    const int64_t v = run_traversal(method, &r, queries + i * D, world_bounds,
                                    (const uint8_t *)tree, n, L, eps);
    ids[i] = r.closest_id;
    if (dist2) dist2[i] = r.closest_d2;
    if (visits) visits[i] = v;
  }
}

/* stackBased/stackFree/cct::knn (knn.h:550-678); stackFree::knn uses
 * eps = 0 (knn.h:659-660). */
void oc_knn(int method, int list_kind, int k, const void *tree, int64_t n,
            const oc_layout *L, const void *world_bounds,
            const float *queries, int64_t nq, float cutoff,
            int32_t *ids, float *dist2, float *ret, int64_t *visits,
            int num_threads)
{
  const int D = L->num_dims;
#pragma omp parallel for schedule(dynamic, 256) num_threads(num_threads > 0 ? num_threads : 1)
  for (int64_t i = 0; i < nq; i++) {
    uint64_t entries[256];
    result_t r;
    r.kind = list_kind == OC_LIST_HEAP ? 2 : 1;
    r.k = k;
    r.id_offset = 0;
    r.entry = entries;
    result_init(&r, cutoff);
    const int64_t v = run_traversal(method, &r, queries + i * D, world_bounds,
                                    (const uint8_t *)tree, n, L, 0.f);
    for (int j = 0; j < k; j++) {
      ids[i * k + j] = (int32_t)(uint32_t)entries[j];
      dist2[i * k + j] = u2f((uint32_t)(entries[j] >> 32));
    }
    if (ret) ret[i] = result_max_r2(&r);
    if (visits) visits[i] = v;
  }
}

/* Multi-round kNN into persistent lists (samples/mpiHugeQuery.cu:75-103):
 * each query's FlexHeapCandidateList (knn.h:266-287, 360-421) lives in
 * lists[i*k .. i*k+k); init != 0 fills it with (cutoff^2, -1) first (the
 * reference's round 0), otherwise the heap is continued as is.  Candidate
 * ids are id_offset + tree index (the reference's rounds use offset 0). */
void oc_knn_accumulate(int method, int k, const void *tree, int64_t n, const oc_layout *L,
                       const void *world_bounds, const float *queries, int64_t nq,
                       uint64_t *lists, int init, float cutoff, int32_t id_offset,
                       int64_t *visits, int num_threads)
{
  const int D = L->num_dims;
#pragma omp parallel for schedule(dynamic, 256) num_threads(num_threads > 0 ? num_threads : 1)
  for (int64_t i = 0; i < nq; i++) {
    result_t r;
    r.kind = 2;
    r.k = k;
    r.id_offset = id_offset;
    r.entry = lists + i * k;
    if (init) result_init(&r, cutoff);
    const int64_t v = run_traversal(method, &r, queries + i * D, world_bounds,
                                    (const uint8_t *)tree, n, L, 0.f);
    if (visits) visits[i] = v;
  }
}

/* extractFinalResult (samples/mpiHugeQuery.cu:90-103): radius = sqrt of
 * the list's max dist2 unless infinite; ids/dist2 (nullable) receive the
 * entries in ascending (dist2, id) order. */
static int cmp_u64(const void *a, const void *b)
{
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}
void oc_knn_extract(const uint64_t *lists, int64_t nq, int k, int32_t *ids, float *dist2,
                    float *radius)
{
  uint64_t tmp[256];
  for (int64_t i = 0; i < nq; i++) {
    const uint64_t *e = lists + i * k;
    float m = u2f((uint32_t)(e[0] >> 32));
    if (radius) radius[i] = isinf(m) ? m : sqrtf(m);
    if (!ids && !dist2) continue;
    memcpy(tmp, e, (size_t)k * 8);
    qsort(tmp, (size_t)k, 8, cmp_u64);
    for (int j = 0; j < k; j++) {
      if (ids) ids[i * k + j] = (int32_t)(uint32_t)tmp[j];
      if (dist2) dist2[i * k + j] = u2f((uint32_t)(tmp[j] >> 32));
    }
  }
}

/* ------------------------------------------------------------------ */
/* spatial k-d tree (spatial-kdtree.h:69-612)                           */
/*                                                                      */
/* Deterministic restatement: nodes are numbered breadth-first, children */
/* of the nodes split in one round allocated in ascending parent id      */
/* order (the reference takes atomicAdd order, spatial-kdtree.h:385);    */
/* degenerate (dim -1) splits send a node's prims alternately left and   */
/* right in ascending primID order (the reference takes atomicAdd        */
/* arrival order, spatial-kdtree.h:424-426).  Leaf dim/pos are written   */
/* as 0/0.0f (the reference leaves union leftovers there).               */
/* ------------------------------------------------------------------ */

/* AtomicBox::encode/decode (spatial-kdtree.h:172-188) */
static inline int32_t abox_enc(float f)
{
  int32_t bits;
  memcpy(&bits, &f, 4);
  if (bits & (int32_t)0x80000000) bits ^= 0x7fffffff;
  return bits;
}
static inline float abox_dec(int32_t bits)
{
  if (bits & (int32_t)0x80000000) bits ^= 0x7fffffff;
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

/* A coordinate as the spatial builder sees it: int coordinates go through
 * the implicit int -> float conversion of AtomicBox::encode(float) and of
 * 'const float center = get_coord(point,split.dim)'
 * (spatial-kdtree.h:172-188, 199-207, 427-429). */
static inline float sp_coord(const uint8_t *e, const oc_layout *L, int d)
{
  return L->scalar_type == OC_F32 ? get_f(e, L, d) : (float)get_i(e, L, d);
}

/* Node::pos and the tree's box_t hold scalar_t: a float result is stored
 * into an int one by C's float -> int conversion (truncation) for int
 * points (spatial-kdtree.h:75-77, 163-168, 489). */
static inline uint32_t sp_store(float f, const oc_layout *L)
{
  uint32_t bits;
  if (L->scalar_type == OC_F32) {
    memcpy(&bits, &f, 4);
  } else {
    const int32_t i = (int32_t)f;
    memcpy(&bits, &i, 4);
  }
  return bits;
}

typedef struct {
  int32_t lo[OC_MAX_DIMS], hi[OC_MAX_DIMS];
  uint32_t count;
  int state; /* 0 open branch, 1 done */
  int dim;
  float pos;
  uint32_t offset;
  uint32_t dcount;
} sp_tmp;

int64_t oc_spatial_build(const void *points, int64_t n, const oc_layout *L,
                         int leaf_threshold, oc_spatial_node *nodes,
                         uint32_t *prim_ids, float *bounds_out)
{
  const int D = L->num_dims;
  if (leaf_threshold == 0) leaf_threshold = 8; /* spatial-kdtree.h:505-506 */
  const uint8_t *base = (const uint8_t *)points;
  int64_t cap = 2 * (n > 0 ? n : 1) + 2;
  sp_tmp *tn = (sp_tmp *)calloc((size_t)cap, sizeof(sp_tmp));
  uint32_t *prim_node = (uint32_t *)malloc((size_t)(n > 0 ? n : 1) * 4);
  uint32_t *split_rank = (uint32_t *)calloc((size_t)cap, 4);
  if (!tn || !prim_node || !split_rank) { free(tn); free(prim_node); free(split_rank); return -1; }

  /* initState (spatial-kdtree.h:276-292) */
  int64_t num_nodes = 2;
  for (int d = 0; d < D; d++) { tn[0].lo[d] = abox_enc(FLT_MAX); tn[0].hi[d] = abox_enc(-FLT_MAX); }
  tn[0].state = 0;
  tn[1].state = 1; tn[1].offset = 0; tn[1].dcount = 0; tn[1].dim = 0; tn[1].pos = 0.f;
  /* initPrims (spatial-kdtree.h:294-314) */
  for (int64_t i = 0; i < n; i++) {
    prim_node[i] = 0;
    tn[0].count++;
    for (int d = 0; d < D; d++) {
      const int32_t v = abox_enc(sp_coord(base + i * L->stride, L, d));
      if (v < tn[0].lo[d]) tn[0].lo[d] = v;
      if (v > tn[0].hi[d]) tn[0].hi[d] = v;
    }
  }
  /* saveBounds (spatial-kdtree.h:464-473): box_t of scalar_t */
  if (bounds_out) {
    uint32_t *bo = (uint32_t *)bounds_out;
    for (int d = 0; d < D; d++) {
      bo[d] = sp_store(abox_dec(tn[0].lo[d]), L);
      bo[D + d] = sp_store(abox_dec(tn[0].hi[d]), L);
    }
  }

  int64_t first_open = 0, end_open = 1;
  while (first_open < end_open) {
    const int64_t round_end = end_open;
    /* selectSplits (spatial-kdtree.h:316-396) for this round's open nodes */
    for (int64_t id = first_open; id < round_end; id++) {
      sp_tmp *t = &tn[id];
      if (t->state == 1) continue;
      if (t->count <= (uint32_t)leaf_threshold) {
        t->state = 1; t->dcount = t->count; t->offset = 0xFFFFFFFFu;
        t->dim = 0; t->pos = 0.f;
        continue;
      }
      float widest_w = 0.f;
      int widest = -1;
      for (int d = 0; d < D; d++) {
        const float w = abox_dec(t->hi[d]) - abox_dec(t->lo[d]);
        if (w <= widest_w) continue;
        widest_w = w;
        widest = d;
      }
      float pos = 0.f;
      if (widest >= 0) {
        pos = 0.5f * (abox_dec(t->lo[widest]) + abox_dec(t->hi[widest]));
        if (pos == abox_dec(t->lo[widest]) || pos == abox_dec(t->hi[widest]))
          widest = -1;
      }
      if (widest < 0) pos = abox_dec(t->lo[0]);
      t->dim = widest;
      t->pos = pos;
      t->offset = (uint32_t)num_nodes;
      t->state = 2; /* split this round */
      for (int s = 0; s < 2; s++) {
        sp_tmp *c = &tn[num_nodes + s];
        for (int d = 0; d < D; d++) { c->lo[d] = abox_enc(FLT_MAX); c->hi[d] = abox_enc(-FLT_MAX); }
        c->count = 0; c->state = 0;
      }
      num_nodes += 2;
    }
    /* updatePrims (spatial-kdtree.h:398-436), prims in ascending id order */
    for (int64_t i = 0; i < n; i++) {
      sp_tmp *t = &tn[prim_node[i]];
      if (t->state != 2) continue;
      const uint8_t *pe = base + i * L->stride;
      int side;
      if (t->dim == -1) side = (int)(split_rank[prim_node[i]]++ & 1u);
      else side = sp_coord(pe, L, t->dim) >= t->pos;
      const uint32_t child = t->offset + (uint32_t)side;
      sp_tmp *ct = &tn[child];
      ct->count++;
      for (int d = 0; d < D; d++) {
        const int32_t v = abox_enc(sp_coord(pe, L, d));
        if (v < ct->lo[d]) ct->lo[d] = v;
        if (v > ct->hi[d]) ct->hi[d] = v;
      }
      prim_node[i] = child;
    }
    for (int64_t id = first_open; id < round_end; id++)
      if (tn[id].state == 2) { tn[id].state = 1; tn[id].dcount = 0; }
    first_open = round_end;
    end_open = num_nodes;
  }

  if (nodes) {
    /* stable sort of prims by node id (CUB radix on bits 32..63,
     * spatial-kdtree.h:577-585), then leaf offsets = first position
     * (writePrimsAndLeafOffsets, spatial-kdtree.h:442-461). */
    int64_t *start = (int64_t *)calloc((size_t)num_nodes + 1, sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) start[prim_node[i] + 1]++;
    for (int64_t j = 0; j < num_nodes; j++) start[j + 1] += start[j];
    int64_t *fill = (int64_t *)malloc(((size_t)num_nodes + 1) * sizeof(int64_t));
    memcpy(fill, start, ((size_t)num_nodes + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) prim_ids[fill[prim_node[i]]++] = (uint32_t)i;
    for (int64_t j = 0; j < num_nodes; j++) {
      sp_tmp *t = &tn[j];
      oc_spatial_node *o = &nodes[j];
      const int is_leaf_with_prims = t->dcount > 0;
      o->offset = is_leaf_with_prims ? (uint32_t)start[j] : t->offset;
      o->count = (uint16_t)t->dcount;
      o->dim = (int16_t)(t->dim == -1 ? 0 : t->dim); /* writeNodes :491-493 */
      const uint32_t pb = sp_store(t->pos, L);
      memcpy(&o->pos, &pb, 4);
    }
    free(start);
    free(fill);
  }
  free(tn); free(prim_node); free(split_rank);
  return num_nodes;
}

/* spatial stackBased::fcp/knn (fcp.h:349-427, knn.h:680-752) and
 * cct::fcp/knn (fcp.h:256-347, knn.h:566-647).  Returns 0, or -1 on
 * stack overflow at depth 50 (the reference prints and returns -1). */
static int spatial_traverse(int method, result_t *res, const float *q,
                            const oc_spatial_node *nodes,
                            const uint32_t *prim_ids, const float *bounds,
                            const uint8_t *data, const oc_layout *L)
{
  enum { DEPTH = 50 };
  const int D = L->num_dims;
  float cull = result_max_r2(res);
  int stack_node[DEPTH];
  float stack_d2[DEPTH];
  float stack_corner[DEPTH][OC_MAX_DIMS];
  int sp = 0;
  int node_id = 0;
  float closest[OC_MAX_DIMS] = {0}, p[OC_MAX_DIMS], far_corner[OC_MAX_DIMS];
  if (method == OC_CCT) {
    for (int d = 0; d < D; d++) {
      float v = q[d];
      const float lo = bounds[d], hi = bounds[D + d];
      v = (v < lo) ? lo : v;
      v = (hi < v) ? hi : v;
      closest[d] = v;
    }
    if (sqr_dist(q, closest, D) > cull) return 0;
  }
  for (;;) {
    oc_spatial_node node;
    for (;;) {
      node = nodes[node_id];
      /* an empty root leaf (N = 0) has count 0 and offset ~0u */
      if (node.count || node.offset == 0xFFFFFFFFu) break;
      const float qc = q[node.dim];
      const int left_close = qc < node.pos;
      const int lchild = (int)node.offset, rchild = (int)node.offset + 1;
      const int close_c = left_close ? lchild : rchild;
      const int far_c = left_close ? rchild : lchild;
      if (method == OC_CCT) {
        memcpy(far_corner, closest, sizeof(float) * D);
        far_corner[node.dim] = node.pos;
        if (sqr_dist(far_corner, q, D) < cull) {
          memcpy(stack_corner[sp], far_corner, sizeof(float) * D);
          stack_node[sp] = far_c;
          ++sp;
          if (sp >= DEPTH) return -1;
        }
      } else {
        const float diff = qc - node.pos;
        const float plane2 = diff * diff;
        if (plane2 < cull) {
          stack_node[sp] = far_c;
          stack_d2[sp] = plane2;
          ++sp;
          if (sp >= DEPTH) return -1;
        }
      }
      node_id = close_c;
    }
    for (int i = 0; i < node.count; i++) {
      const uint32_t prim = prim_ids[node.offset + (uint32_t)i];
      load_point(data + (size_t)prim * L->stride, L, p);
      cull = result_process(res, (int)prim, sqr_dist(p, q, D));
    }
    for (;;) {
      if (sp == 0) return 0;
      --sp;
      if (method == OC_CCT) {
        memcpy(closest, stack_corner[sp], sizeof(float) * D);
        if (sqr_dist(closest, q, D) >= cull) continue;
      } else if (stack_d2[sp] >= cull) {
        continue;
      }
      node_id = stack_node[sp];
      break;
    }
  }
}

void oc_spatial_fcp(int method, const oc_spatial_node *nodes,
                    const uint32_t *prim_ids, const float *bounds,
                    const void *points, const oc_layout *L,
                    const float *queries, int64_t nq, float cutoff,
                    int32_t *ids, float *dist2, int num_threads)
{
  const int D = L->num_dims;
#pragma omp parallel for schedule(dynamic, 256) num_threads(num_threads > 0 ? num_threads : 1)
  for (int64_t i = 0; i < nq; i++) {
    result_t r;
    r.kind = 0; r.k = 1; r.entry = NULL; r.id_offset = 0;
    result_init(&r, cutoff);
    const int st = spatial_traverse(method, &r, queries + i * D, nodes, prim_ids,
                                    bounds, (const uint8_t *)points, L);
    ids[i] = st < 0 ? -1 : r.closest_id;
    if (dist2) dist2[i] = r.closest_d2;
  }
}

void oc_spatial_knn(int method, int list_kind, int k,
                    const oc_spatial_node *nodes, const uint32_t *prim_ids,
                    const float *bounds, const void *points,
                    const oc_layout *L, const float *queries, int64_t nq,
                    float cutoff, int32_t *ids, float *dist2, float *ret,
                    int num_threads)
{
  const int D = L->num_dims;
#pragma omp parallel for schedule(dynamic, 256) num_threads(num_threads > 0 ? num_threads : 1)
  for (int64_t i = 0; i < nq; i++) {
    uint64_t entries[256];
    result_t r;
    r.kind = list_kind == OC_LIST_HEAP ? 2 : 1;
    r.k = k;
    r.id_offset = 0;
    r.entry = entries;
    result_init(&r, cutoff);
    const int st = spatial_traverse(method, &r, queries + i * D, nodes, prim_ids,
                                    bounds, (const uint8_t *)points, L);
    for (int j = 0; j < k; j++) {
      ids[i * k + j] = (int32_t)(uint32_t)entries[j];
      dist2[i * k + j] = u2f((uint32_t)(entries[j] >> 32));
    }
    if (ret) ret[i] = st < 0 ? -1.f : result_max_r2(&r);
  }
}

/* computeHash (testing/testBuildersSameResult.cu:75-82) */
uint64_t oc_hash_words(const void *ptr, int64_t num_bytes)
{
  uint64_t h = 0;
  const uint32_t *w = (const uint32_t *)ptr;
  for (int64_t i = 0; i < num_bytes / 4; i++) h = h * 17 ^ w[i];
  return h;
}
