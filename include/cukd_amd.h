/* cukd_amd.h -- C ABI of the gfx950 (MI355X) k-d tree library.
 *
 * This is the drop-in boundary for cudaKDTree's build-and-query path.  Each
 * entry point names the reference interface it replaces (file:line in the
 * ingowald/cudaKDTree v1.0.1 tree).  The C++ header API under include/cukd/
 * maps the reference's templates onto these calls; other languages bind
 * them directly (cudakdtree_amd/__init__.py is a ctypes binding).
 *
 * Conventions
 *  - All pointers named d_* are device pointers.  Every call is
 *    stream-ordered on 'stream' and returns before the GPU work is done,
 *    except where noted (the NaN screen of cukd_build_tree, the host
 *    builder, spatial builds, forest set-up and tuning runs wait).
 *  - Return value: CUKD_OK (0) or a negative CUKD_ERR_* code; the message
 *    of the last error on the calling thread is in cukd_last_error().
 *    Nothing throws across this boundary.
 *  - Sizes are int32 like the reference's 'int numPoints' (N < 2^30,
 *    the depth limit of the reference's 30-entry traversal stacks).
 */
#ifndef CUKD_AMD_H
#define CUKD_AMD_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUKD_OK 0
#define CUKD_ERR_INVALID (-1)     /* bad argument (null pointer, bad layout) */
#define CUKD_ERR_HIP (-2)         /* a HIP runtime call failed               */
#define CUKD_ERR_NO_BOUNDS (-3)   /* explicit dims but no world-bounds buffer */
#define CUKD_ERR_ALLOC (-4)       /* the memory resource failed              */
#define CUKD_ERR_UNSUPPORTED (-5) /* combination not implemented            */

/* Runtime description of one data_t element, replacing the compile-time
 * data_traits/point_traits pair (data.h:56-116, cukd-math.h:298-331):
 * num_dims scalar coordinates stored contiguously at point_offset, an
 * optional split-dim field (data_traits::has_explicit_dim, get_dim/set_dim,
 * data.h:98-113) and the element stride.  Covers float2/3/4, int2/3/4,
 * vec_float<N>, and payload structs such as PointPlusPayload or Photon. */
typedef struct cukd_layout {
  int32_t num_dims;     /* 1..8                                         */
  int32_t scalar_type;  /* CUKD_SCALAR_F32 or CUKD_SCALAR_I32           */
  int32_t stride;       /* sizeof(data_t), a multiple of 4, <= 256      */
  int32_t point_offset; /* byte offset of coordinate 0, multiple of 4   */
  int32_t dim_offset;   /* byte offset of the split-dim field, -1: none */
  int32_t dim_bytes;    /* 1, 2 or 4                                    */
} cukd_layout;

#define CUKD_SCALAR_F32 0
#define CUKD_SCALAR_I32 1

/* Allocation policy for builder temporaries, replacing
 * cukd::GpuMemoryResource (helpers.h:31-34).  NULL selects the default,
 * stream-ordered hipMallocAsync/hipFreeAsync (AsyncGpuMemoryResource,
 * helpers.h:52-60).  Callbacks return 0 on success. */
typedef struct cukd_mem_resource {
  void *ctx;
  int (*malloc_fn)(void *ctx, void **ptr, size_t bytes, hipStream_t s);
  int (*free_fn)(void *ctx, void *ptr, hipStream_t s);
} cukd_mem_resource;

/* builder selection (builder.h:108-151) */
#define CUKD_BUILDER_ID_DEFAULT 0 /* cukd::buildTree                        */
#define CUKD_BUILDER_ID_THRUST 1  /* cukd::buildTree_thrust                 */
#define CUKD_BUILDER_ID_BITONIC 2 /* cukd::buildTree_bitonic                */
#define CUKD_BUILDER_ID_INPLACE 3 /* cukd::buildTree_inPlace                */

/* traversal selection (fcp.h / knn.h namespaces) */
#define CUKD_STACK_BASED 0 /* cukd::stackBased:: (traverse-default-stack-based.h) */
#define CUKD_STACK_FREE 1  /* cukd::stackFree::  (traverse-stack-free.h)          */
#define CUKD_CCT 2         /* cukd::cct::        (traverse-cct.h)                 */

/* candidate lists (knn.h:214-258) */
#define CUKD_LIST_FIXED 0 /* FixedCandidateList<k>: sorted ascending */
#define CUKD_LIST_HEAP 1  /* HeapCandidateList<k>: max-heap order    */

/* query flags */
#define CUKD_QUERY_SORT 1 /* Morton-order the queries before traversal
                             (results are returned in input order)      */
#define CUKD_QUERY_AUTOTUNE 2 /* first call per (dims, traversal, list, k,
                             log2 nq): time workgroup sizes 64/128/256 with
                             and without Morton ordering, and one or two
                             queries per lane where two apply, on the caller's
                             inputs (synchronises the stream once), keep
                             the fastest for later calls; results are the
                             same for every choice.  Not for multi-round
                             lists (they only use an already tuned choice) */
#define CUKD_QUERY_ONE_PER_LANE 4 /* one query per lane everywhere.  By
                             default the stack-based traversal with FCP or
                             a FixedCandidateList of k = 1/4/8/16 over a
                             tree with round-robin split dims (and under
                             4 GiB) steps two queries per lane, so twice as
                             many node loads are in flight; results are
                             identical either way */
#define CUKD_QUERY_PRESORTED 8 /* the queries are already in Morton order
                             (e.g. permuted by cukd_query_order): launch
                             with the shape of a CUKD_QUERY_SORT launch,
                             without ordering them again */

/* Builds a left-balanced k-d tree in place over d_points[0..n).
 * Replaces template<data_t,traits> buildTree / buildTree_thrust /
 * buildTree_bitonic / buildTree_inPlace (builder.h:108-122,
 * builder_thrust.h:90-110, builder_bitonic.h:83-94,
 * builder_inplace.h:82-102).  d_world_bounds (box of 2*num_dims scalars,
 * lower then upper) receives the bounds when non-null and is required for
 * explicit-dim layouts (builder_thrust.h:469-473).  n < 1 returns CUKD_OK
 * without work (builder_thrust.h:441).
 *
 * CUKD_BUILDER_ID_DEFAULT / _THRUST: radix-select builder, about 2.3n words
 *   of temporary memory; the array is identical to buildTree_host's
 *   (builder_thrust.h:514-601) for every input without NaN coordinates,
 *   ties included.
 * CUKD_BUILDER_ID_BITONIC: sorts each subtree inside the final array with a
 *   bitonic network; temporary memory n ints (builder.h:34-38), identical
 *   to buildTree_host's like the default builder.
 * CUKD_BUILDER_ID_INPLACE: same network, no temporary memory at all
 *   (builder.h:41-45); ties on a split coordinate are broken by the
 *   element's bytes, so the array equals buildTree_host's whenever the
 *   split coordinates decide (always on tie-free inputs) and is otherwise
 *   a valid, deterministic left-balanced tree.
 *
 * NaN coordinates: the reference compares coordinates with '<'
 * (builder_thrust.h:401-404), which orders nothing once a NaN is present.
 * For float layouts every builder first screens the coordinates (one read
 * of the points and one wait for 'stream', as the reference's thrust
 * builder waits once per level, builder_thrust.h:500) and returns CUKD_ERR_INVALID,
 * with the array unchanged, if any is NaN.  Under HIP stream capture the
 * screen is skipped, so the call stays capturable and does not wait. */
int cukd_build_tree(void *d_points, int32_t n, const cukd_layout *layout,
                    int builder, void *d_world_bounds, hipStream_t stream,
                    const cukd_mem_resource *mem);

/* buildTree_host (builder_thrust.h:114-118, 514-601): the same tree over
 * points in host (or managed) memory, h_world_bounds (host, nullable unless
 * explicit dims) receiving the bounds.  The points go to the GPU, are built
 * by the default builder and come back; the call returns when the array is
 * final.  Errors as cukd_build_tree. */
int cukd_build_tree_host(void *h_points, int32_t n, const cukd_layout *layout,
                         void *h_world_bounds);

/* Bytes of temporary device memory cukd_build_tree will request. */
int cukd_build_tree_temp_bytes(int32_t n, const cukd_layout *layout,
                               int builder, size_t *bytes);

/* World bounds of d_points (builder_common.h:59-64, 151-163). */
int cukd_compute_bounds(void *d_bounds, const void *d_points, int32_t n,
                        const cukd_layout *layout, hipStream_t stream);

/* Batched find-closest-point over a built tree: one query per d_queries
 * row (num_dims floats).  Replaces per-thread calls of
 * stackBased/stackFree/cct::fcp (fcp.h:61-163).  d_ids[i] is the tree
 * index of the closest point with dist2 < cutoff^2, or -1; d_dist2
 * (nullable) receives that dist2 (cutoff^2 when none).  eps is
 * FcpSearchParams::eps (stack-free only, fcp.h:237).  d_world_bounds is
 * required for CUKD_CCT. */
int cukd_fcp(int traversal, const void *d_tree, int32_t n,
             const cukd_layout *layout, const void *d_world_bounds,
             const float *d_queries, int32_t nq, float cutoff, float eps,
             int32_t *d_ids, float *d_dist2, int flags, hipStream_t stream,
             const cukd_mem_resource *mem);

/* Batched k-nearest-neighbour query.  Replaces per-thread calls of
 * stackBased/stackFree/cct::knn with a FixedCandidateList<k> or
 * HeapCandidateList<k> (knn.h:83-199, 214-258).  For query i the k list
 * entries are written to d_ids[i*k..] (int(uint32(entry)), -1 = empty) and
 * d_dist2[i*k..], in the list's own storage order; d_max_radius2[i]
 * (nullable) receives the call's return value, maxRadius2().  1 <= k <= 64. */
int cukd_knn(int traversal, int list_kind, int k, const void *d_tree,
             int32_t n, const cukd_layout *layout,
             const void *d_world_bounds, const float *d_queries, int32_t nq,
             float cutoff, int32_t *d_ids, float *d_dist2,
             float *d_max_radius2, int flags, hipStream_t stream,
             const cukd_mem_resource *mem);

/* cukd_knn without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors as cukd_knn, less the memory resource.
 * CUKD_QUERY_SORT orders the queries in a library-owned device array of
 * 52 MiB, one per GPU, allocated when the library's code is loaded and held
 * until the process ends, so nq is limited to 4194304 (2^22; larger is
 * CUKD_ERR_INVALID).  CUKD_QUERY_AUTOTUNE only uses launch shapes an
 * earlier cukd_knn call has tuned; it never tunes.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (several streams or
 * threads at once) race on it and must not be made. */
int cukd_knn_noalloc(int traversal, int list_kind, int k, const void *d_tree,
                     int32_t n, const cukd_layout *layout,
                     const void *d_world_bounds, const float *d_queries,
                     int32_t nq, float cutoff, int32_t *d_ids, float *d_dist2,
                     float *d_max_radius2, int flags, hipStream_t stream);

/* The launch order CUKD_QUERY_SORT uses: d_order (nq int32) receives a
 * permutation of 0..nq-1 listing the queries by the Morton code of their
 * coordinates (15-bit codes over the batch's bounding box; the order of
 * queries inside one code cell is unspecified).  No counterpart in the
 * reference. */
int cukd_query_order(const float *d_queries, int32_t nq, int32_t num_dims,
                     int32_t *d_order, hipStream_t stream,
                     const cukd_mem_resource *mem);

/* One round of a multi-round kNN: every query continues its own
 * FlexHeapCandidateList (knn.h:266-287, 360-421) held in HBM at
 * d_lists[q*k .. q*k+k) (uint64 entries, dist2 bits << 32 | id), as the
 * reference's runQuery kernel does per ring round
 * (samples/mpiHugeQuery.cu:75-88).  init != 0 first fills the list with
 * (cutoff^2, -1) (the reference's round 0); otherwise the list is used as
 * it is.  Candidate ids are id_offset + tree index (the reference's rounds
 * use 0, leaving ids local to each tree).  k in 1..64. */
int cukd_knn_accumulate(int traversal, int k, const void *d_tree, int32_t n,
                        const cukd_layout *layout, const void *d_world_bounds,
                        const float *d_queries, int32_t nq, uint64_t *d_lists,
                        int init, float cutoff, int32_t id_offset, int flags,
                        hipStream_t stream, const cukd_mem_resource *mem);

/* extractFinalResult (samples/mpiHugeQuery.cu:90-103): d_radius[q] =
 * sqrt of the list's max dist2 (infinity stays infinity).  d_ids/d_dist2
 * (each nullable, nq*k) receive the entries in ascending (dist2, id)
 * order. */
int cukd_knn_extract(const uint64_t *d_lists, int32_t nq, int k,
                     int32_t *d_ids, float *d_dist2, float *d_radius,
                     hipStream_t stream);

/* Spatial k-d tree (spatial-kdtree.h:69-98).  Node = {scalar_t pos;
 * uint32 offset; uint16 count; int16 dim} (12 bytes): pos and the bounds
 * are float for float points and int32 for int points (ipos, ibounds). */
typedef struct cukd_spatial_node {
  union {
    float pos;
    int32_t ipos;
  };
  uint32_t offset;
  uint16_t count;
  int16_t dim;
} cukd_spatial_node;

typedef struct cukd_spatial_tree {
  union {                        /* lower[num_dims] then upper[num_dims] */
    float bounds[2 * 8];
    int32_t ibounds[2 * 8];
  };
  cukd_spatial_node *nodes;      /* device, num_nodes entries            */
  uint32_t *prim_ids;            /* device, num_prims entries            */
  const void *data;              /* device, the (unpermuted) points      */
  int32_t num_prims;
  int32_t num_nodes;
} cukd_spatial_tree;

/* Builds a spatial k-d tree (spatial-kdtree.h:113-120, 496-612, 632-645).
 * leaf_threshold 0 means 8 (BuildConfig::makeLeafThreshold, :100-105).
 * Int points are built as the reference builds them: the split decisions
 * see each coordinate converted to float, and pos and the bounds are that
 * float value converted back to int32 (truncated).  Queries need float
 * points (the reference has no distance math for int points).
 * Synchronises 'stream' before returning, as the reference does, because
 * the node count and bounds are returned in host memory. */
int cukd_spatial_build(cukd_spatial_tree *tree, const void *d_points,
                       int32_t n, const cukd_layout *layout,
                       int leaf_threshold, hipStream_t stream,
                       const cukd_mem_resource *mem);

/* Releases a spatial tree's nodes/primIDs (spatial-kdtree.h:618-630). */
int cukd_spatial_free(cukd_spatial_tree *tree, hipStream_t stream,
                      const cukd_mem_resource *mem);

/* Spatial FCP / kNN (fcp.h:256-427, knn.h:566-752); traversal is
 * CUKD_STACK_BASED or CUKD_CCT.  Ids are original point indices. */
int cukd_spatial_fcp(int traversal, const cukd_spatial_tree *tree,
                     const cukd_layout *layout, const float *d_queries,
                     int32_t nq, float cutoff, int32_t *d_ids, float *d_dist2,
                     hipStream_t stream);
int cukd_spatial_knn(int traversal, int list_kind, int k,
                     const cukd_spatial_tree *tree, const cukd_layout *layout,
                     const float *d_queries, int32_t nq, float cutoff,
                     int32_t *d_ids, float *d_dist2, float *d_max_radius2,
                     hipStream_t stream);

/* ---- forests: many independent trees, built and queried together ----
 *
 * No counterpart in the reference, which builds and queries one tree per
 * call; every tree of a forest is the tree cukd_build_tree builds over the
 * same points, and every query result equals cukd_fcp / cukd_knn on that
 * tree alone (ids index the query's own tree). */
typedef struct cukd_forest cukd_forest;

/* Plans a forest: tree t is d_points[h_offsets[t] .. h_offsets[t+1]) of
 * one array (h_offsets: host, num_trees+1 values, starting at 0,
 * non-decreasing, every tree < 2^30 points, all < 2^31).  Allocates the
 * whole build workspace once through mem (NULL: hipMallocAsync) and
 * synchronises stream once, so that cukd_forest_build allocates nothing,
 * never synchronises, and can be captured into a HIP graph. */
int cukd_forest_create(cukd_forest **forest, const int64_t *h_offsets,
                       int32_t num_trees, const cukd_layout *layout,
                       hipStream_t stream, const cukd_mem_resource *mem);

/* Device bytes held by the forest's workspace. */
int cukd_forest_workspace_bytes(const cukd_forest *forest, size_t *bytes);

/* Builds every tree of the forest in place, in one sequence of launches
 * for all trees.  d_world_bounds (nullable unless explicit dims) receives
 * one box per tree, num_trees * 2 * num_dims scalars. */
int cukd_forest_build(cukd_forest *forest, void *d_points,
                      void *d_world_bounds, hipStream_t stream);

int cukd_forest_destroy(cukd_forest *forest, hipStream_t stream);

/* Batched queries over a built forest in one launch.  The nq queries form
 * num_segments segments: segment s is queries
 * d_seg_offsets[s] .. d_seg_offsets[s+1]) and runs against tree
 * d_seg_tree[s] (both device arrays; a tree may appear in several
 * segments).  d_world_bounds (per-tree boxes) is required for CUKD_CCT.
 * Output layout and meaning as cukd_fcp / cukd_knn; flags as there
 * (CUKD_QUERY_SORT orders the whole batch). */
int cukd_forest_fcp(int traversal, const cukd_forest *forest,
                    const void *d_points, const void *d_world_bounds,
                    const float *d_queries, int32_t nq, int32_t num_segments,
                    const int32_t *d_seg_offsets, const int32_t *d_seg_tree,
                    float cutoff, float eps, int32_t *d_ids, float *d_dist2,
                    int flags, hipStream_t stream);
int cukd_forest_knn(int traversal, int list_kind, int k,
                    const cukd_forest *forest, const void *d_points,
                    const void *d_world_bounds, const float *d_queries,
                    int32_t nq, int32_t num_segments,
                    const int32_t *d_seg_offsets, const int32_t *d_seg_tree,
                    float cutoff, int32_t *d_ids, float *d_dist2,
                    float *d_max_radius2, int flags, hipStream_t stream);

/* Reserves the Morton-ordering scratch for forest queries of up to
 * max_queries queries on 'stream', so that CUKD_QUERY_SORT forest queries
 * on that stream allocate nothing (e.g. inside a captured HIP graph).  The
 * scratch serves that one stream: sorted forest queries on any other
 * stream allocate scratch of their own (through the forest's memory
 * resource), so queries on several streams never share it.  Reserving
 * again from another stream moves the scratch to that stream. */
int cukd_forest_reserve_queries(cukd_forest *forest, int32_t max_queries,
                                hipStream_t stream);

/* ---- serving loop over a resident forest ----
 *
 * No counterpart in the reference.  Requests (one tree, nq host queries,
 * FCP or kNN-k) are submitted from any thread; a worker thread packs
 * whatever is queued into one batch slot (admitted in arrival order while
 * the batch, laid out exactly as it will be launched, fits one eighth of
 * budget_bytes of device memory, at most max_requests_per_batch requests,
 * <= 0 for no limit) and launches one forest query per (k, cutoff) class.
 * Inside a batch, each (k, cutoff) class is launched as a light group and
 * a heavy group (nq * max(k, 1) >= 65536), light groups first, and each
 * group's requests complete as soon as their own launch has run, so a
 * light request never waits for a heavy one.  A request too large for one
 * slot is queued in pieces that each fit one and completes with its last
 * piece.  Up to eight batches are in flight, each on a stream of its own,
 * and each is completed (results copied back, requests completed) as soon
 * as it finishes, so a batch slowed by a few long traversals does not
 * delay the batches launched after it.  max_requests_per_batch = 1 serves
 * each request with its own launches.  Results equal cukd_fcp / cukd_knn
 * (fixed list) on the request's tree alone.  A failed batch fails its own
 * requests (their completion status); only a failed HIP call makes later
 * polls report an error. */
typedef struct cukd_server cukd_server;

typedef struct cukd_completion {
  uint64_t ticket;
  int32_t status;       /* CUKD_OK or an error code */
  int32_t num_queries;
  float queue_ms;       /* submit -> admitted into a batch */
  float total_ms;       /* submit -> results in the caller's buffers */
} cukd_completion;

typedef struct cukd_server_stats {
  uint64_t requests_submitted;
  uint64_t requests_done;
  uint64_t queries_done;
  uint64_t batches;
} cukd_server_stats;

/* d_points / d_world_bounds: the built forest's arrays (bounds required
 * for CUKD_CCT).  Allocates eight device batch buffers and their pinned host
 * mirrors (budget_bytes in total), creates the batches' streams, which
 * start after the work queued on 'stream' so far (the forest build, say),
 * and starts the worker thread. */
int cukd_server_create(cukd_server **server, const cukd_forest *forest,
                       const void *d_points, const void *d_world_bounds,
                       int traversal, size_t budget_bytes,
                       int32_t max_requests_per_batch, hipStream_t stream);

/* Queues a request.  k = 0: FCP (h_ids[nq], h_dist2[nq] nullable);
 * k in 1..64: kNN (h_ids / h_dist2 nq*k, ascending).  h_queries (nq rows
 * of num_dims floats) and the output buffers must stay valid until the
 * request's completion is returned by cukd_server_poll. */
int cukd_server_submit(cukd_server *server, int32_t tree,
                       const float *h_queries, int32_t nq, int k, float cutoff,
                       int32_t *h_ids, float *h_dist2, uint64_t *ticket);

/* Returns up to max completions (first waiting until at least wait_for
 * are available, if wait_for > 0). */
int cukd_server_poll(cukd_server *server, cukd_completion *out, int32_t max,
                     int32_t *count, int32_t wait_for);

int cukd_server_get_stats(cukd_server *server, cukd_server_stats *stats);

/* Serves every queued request, stops the worker and frees the buffers. */
int cukd_server_destroy(cukd_server *server);

/* Traversal statistics, replacing CUKD_STATS / cukd::g_traversalStats
 * (common.h:348-360): while d_counter is non-null, every query launch adds
 * the number of tree nodes its traversals visited (one per visited node,
 * as at traverse-default-stack-based.h:64) to *d_counter.  NULL disables. */
int cukd_set_traversal_stats(unsigned long long *d_counter);

/* Message of the last failed call on this thread ("" if none). */
const char *cukd_last_error(void);

/* Library version, 100*major + minor. */
int cukd_version(void);

#ifdef __cplusplus
}
#endif
#endif
