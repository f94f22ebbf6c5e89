/* cukd_oracle.h -- CPU restatement of cudaKDTree's build-and-query path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle: a plain-C
 * restatement of the reference algorithms, used by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py.  The product library
 * (cudakdtree_amd/) never links, loads or calls anything in this folder.
 *
 * Every function names the reference file:line it restates.  Citations
 * are relative to the ingowald/cudaKDTree v1.0.1 source tree.
 */
#ifndef CUKD_ORACLE_H
#define CUKD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Describes one data_t element: D scalar coordinates stored contiguously
 * at point_offset, an optional split-dim field, and the element stride.
 * This is the runtime form of cukd::default_data_traits (data.h:56-116). */
typedef struct oc_layout {
  int32_t num_dims;     /* D, 1..8                                   */
  int32_t scalar_type;  /* 0 = float32, 1 = int32                    */
  int32_t stride;       /* sizeof(data_t), multiple of 4             */
  int32_t point_offset; /* byte offset of coordinate 0               */
  int32_t dim_offset;   /* byte offset of split-dim field, -1 = none */
  int32_t dim_bytes;    /* 1, 2 or 4                                  */
} oc_layout;

enum { OC_F32 = 0, OC_I32 = 1 };
enum { OC_STACK_BASED = 0, OC_STACK_FREE = 1, OC_CCT = 2 };
enum { OC_LIST_FIXED = 0, OC_LIST_HEAP = 1 };

/* ---- tree index math (helpers.h:76-258) ---- */
int     oc_level_of(int64_t node);
int     oc_num_levels_for(int64_t n);
int64_t oc_subtree_size(int64_t node, int64_t n);
int64_t oc_segment_begin(int step, int64_t n, int64_t subtree);
int64_t oc_pivot_pos(int step, int64_t n, int64_t subtree);

/* ---- bounds + widest dim (builder_common.h:166-174, box.h:110-125) ---- */
void oc_compute_bounds(const void *points, int64_t n, const oc_layout *L,
                       void *box_out);
int  oc_widest_dim(const void *box, const oc_layout *L);

/* ---- builder (builder_thrust.h:514-601) ----
 * Returns 0 on success, -1 if explicit dims are requested without a
 * world-bounds buffer (builder_thrust.h:566-570), -2 on allocation error. */
int oc_build_host(void *points, int64_t n, const oc_layout *L,
                  void *world_bounds);

/* ---- balanced-tree queries (fcp.h, knn.h, traverse-*.h) ----
 * queries: nq * D floats.  world_bounds: box of 2*D floats (cct only).
 * ids/dist2 are written per query (FCP) or per query*k (kNN).
 * visits (nullable) receives per-query node-visit counts. */
void oc_fcp(int method, const void *tree, int64_t n, const oc_layout *L,
            const void *world_bounds, const float *queries, int64_t nq,
            float cutoff, float eps, int32_t *ids, float *dist2,
            int64_t *visits, int num_threads);
void oc_knn(int method, int list_kind, int k, const void *tree, int64_t n,
            const oc_layout *L, const void *world_bounds,
            const float *queries, int64_t nq, float cutoff,
            int32_t *ids, float *dist2, float *ret, int64_t *visits,
            int num_threads);

/* multi-round kNN into persistent FlexHeap lists and their extraction
 * (knn.h:266-287, 360-421; samples/mpiHugeQuery.cu:75-103) */
void oc_knn_accumulate(int method, int k, const void *tree, int64_t n, const oc_layout *L,
                       const void *world_bounds, const float *queries, int64_t nq,
                       uint64_t *lists, int init, float cutoff, int32_t id_offset,
                       int64_t *visits, int num_threads);
void oc_knn_extract(const uint64_t *lists, int64_t nq, int k, int32_t *ids, float *dist2,
                    float *radius);

/* ---- spatial k-d tree (spatial-kdtree.h:69-105, 496-612) ---- */
/* pos (and bounds_out) hold scalar_t bits: float, or int32 for int points */
typedef struct oc_spatial_node {
  float    pos;
  uint32_t offset;
  uint16_t count;
  int16_t  dim;
} oc_spatial_node;

/* Two-call protocol: first with nodes==NULL to get the node count, then
 * with buffers of num_nodes nodes and n primIDs.  bounds_out = 2*D floats. */
int64_t oc_spatial_build(const void *points, int64_t n, const oc_layout *L,
                         int leaf_threshold, oc_spatial_node *nodes,
                         uint32_t *prim_ids, float *bounds_out);
void oc_spatial_fcp(int method, const oc_spatial_node *nodes,
                    const uint32_t *prim_ids, const float *bounds,
                    const void *points, const oc_layout *L,
                    const float *queries, int64_t nq, float cutoff,
                    int32_t *ids, float *dist2, int num_threads);
void oc_spatial_knn(int method, int list_kind, int k,
                    const oc_spatial_node *nodes, const uint32_t *prim_ids,
                    const float *bounds, const void *points,
                    const oc_layout *L, const float *queries, int64_t nq,
                    float cutoff, int32_t *ids, float *dist2, float *ret,
                    int num_threads);

/* FNV-like hash used by testing/testBuildersSameResult.cu:75-82. */
uint64_t oc_hash_words(const void *ptr, int64_t num_bytes);

#ifdef __cplusplus
}
#endif
#endif
