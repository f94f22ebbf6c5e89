// plain_forest.cpp -- the per-tree baseline of the batched forest bench,
// timed from C++: for every tree, one cukd_build_tree and one cukd_knn
// through the C ABI, as a caller of the single-tree API (the reference's
// buildTree + a knn launch) would run them.  bench.py loads this library
// and calls cukd_plain_forest_run; keeping the loop in C++ leaves no
// interpreter overhead in the baseline.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>

#include "cukd_amd.h"

extern "C" {

// Restores d_points from d_orig, then builds and queries every tree in
// turn on 'stream'; repeated reps times after one untimed run.  *ms_out =
// mean wall time of one pass (host clock around the timed passes, the
// stream synchronised before and after).  Returns CUKD_OK or the first
// error code.
int cukd_plain_forest_run(const float *d_orig, float *d_points, const int64_t *h_off,
                          int32_t num_trees, const float *d_queries, const int64_t *h_qoff,
                          int32_t k, int32_t *d_ids, float *d_dist2, float *d_radius2,
                          int32_t flags, hipStream_t stream, int32_t reps, double *ms_out)
{
  const cukd_layout L = {3, CUKD_SCALAR_F32, 12, 0, -1, 4};
  const size_t total = (size_t)h_off[num_trees] * 12;
  auto pass = [&]() -> int {
    if (hipMemcpyAsync(d_points, d_orig, total, hipMemcpyDeviceToDevice, stream) != hipSuccess)
      return CUKD_ERR_HIP;
    for (int32_t t = 0; t < num_trees; t++) {
      const int64_t a = h_off[t], n = h_off[t + 1] - a;
      const int64_t qa = h_qoff[t], nq = h_qoff[t + 1] - qa;
      if (n < 1) continue;
      int rc = cukd_build_tree(d_points + 3 * a, (int32_t)n, &L, CUKD_BUILDER_ID_DEFAULT, nullptr,
                               stream, nullptr);
      if (rc != CUKD_OK) return rc;
      if (nq < 1) continue;
      rc = cukd_knn(CUKD_STACK_BASED, CUKD_LIST_FIXED, k, d_points + 3 * a, (int32_t)n, &L,
                    nullptr, d_queries + 3 * qa, (int32_t)nq, INFINITY, d_ids + (size_t)k * qa,
                    d_dist2 + (size_t)k * qa, d_radius2 + qa, flags, stream, nullptr);
      if (rc != CUKD_OK) return rc;
    }
    return CUKD_OK;
  };
  int rc = pass();
  if (rc != CUKD_OK) return rc;
  if (hipStreamSynchronize(stream) != hipSuccess) return CUKD_ERR_HIP;
  const auto t0 = std::chrono::steady_clock::now();
  for (int32_t r = 0; r < reps && rc == CUKD_OK; r++) rc = pass();
  if (hipStreamSynchronize(stream) != hipSuccess) return CUKD_ERR_HIP;
  const auto t1 = std::chrono::steady_clock::now();
  if (ms_out)
    *ms_out = std::chrono::duration<double, std::milli>(t1 - t0).count() / (reps > 0 ? reps : 1);
  return rc;
}

}  // extern "C"
