// Compiles the way the reference's own callers include the library
// (testing/testHostBuilder*.cu, testBuildersSameResult.cu,
// floatN-knn-and-fcp.cu: builder_host.h, builder_thrust.h,
// builder_bitonic.h, builder_inplace.h, the traverse-*.h headers and
// common.h with CUKD_STATS), and checks on the GPU:
//  1. buildTree_host on host memory equals buildTree_thrust on the device,
//     bounds included (builder_thrust.h:114-118, 514-601);
//  2. buildTree picks its builder at compile time (builder.h:124-150):
//     counted through a memory resource, CUKD_BUILDER_INPLACE allocates
//     nothing, CUKD_BUILDER_BITONIC allocates exactly n ints, and the
//     default (thrust) builder allocates its workspace; the tree equals the
//     host builder's on tie-free input;
//  3. with CUKD_ENABLE_STATS, cukd::g_traversalStats counts the node
//     visits of device-side traversals in a user kernel, equal to the
//     library's own batch count (common.h:345-360);
//  4. a NaN coordinate makes buildTree throw and leaves the array as it
//     was.
// Built three times (tests/cpp/Makefile): default, -DCUKD_BUILDER_BITONIC
// and -DCUKD_BUILDER_INPLACE, each with -DCUKD_ENABLE_STATS=1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "cukd/builder_host.h"
#include "cukd/builder_thrust.h"
#include "cukd/builder_bitonic.h"
#include "cukd/builder_inplace.h"
#include "cukd/builder.h"
#include "cukd/common.h"
#include "cukd/traverse-default-stack-based.h"
#include "cukd/traverse-stack-free.h"
#include "cukd/traverse-cct.h"
#include "cukd/fcp.h"
#include "cukd/knn.h"

#define CHECK(x)                                                                              \
  do {                                                                                        \
    hipError_t e = (x);                                                                       \
    if (e != hipSuccess) {                                                                    \
      printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__);                        \
      return 1;                                                                               \
    }                                                                                         \
  } while (0)

// counts what the builder asks of its memory resource
struct CountingResource : cukd::GpuMemoryResource {
  int mallocs = 0, frees = 0;
  size_t bytes = 0;
  hipError_t malloc(void **p, size_t size, hipStream_t s) override
  {
    mallocs++;
    bytes += size;
    return hipMallocAsync(p, size, s);
  }
  hipError_t free(void *p, hipStream_t s) override
  {
    frees++;
    return hipFreeAsync(p, s);
  }
};

__global__ void userQueries(const float3 *tree, int n, cukd::box_t<float3> wb, const float3 *q,
                            int nq, int method)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  if (method == CUKD_STACK_BASED) {
    cukd::FixedCandidateList<8> l(INFINITY);
    cukd::stackBased::knn(l, q[i], tree, n);
  } else if (method == CUKD_STACK_FREE) {
    cukd::FixedCandidateList<8> l(INFINITY);
    cukd::stackFree::knn(l, q[i], tree, n);
  } else {
    cukd::FixedCandidateList<8> l(INFINITY);
    cukd::cct::knn(l, q[i], wb, tree, n);
  }
}

static uint64_t sm(uint64_t &x)
{
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int main()
{
#if defined(CUKD_BUILDER_INPLACE)
  const char *variant = "inplace";
#elif defined(CUKD_BUILDER_BITONIC)
  const char *variant = "bitonic";
#else
  const char *variant = "thrust";
#endif
  const int n = 200000, nq = 2048;
  // distinct coordinates per axis (a shuffled grid of 2^-20 steps), so that
  // every builder must give the host builder's tree
  std::vector<float3> pts(n), qs(nq);
  {
    uint64_t s = 7;
    std::vector<int> perm[3];
    for (int d = 0; d < 3; d++) {
      perm[d].resize(n);
      for (int i = 0; i < n; i++) perm[d][i] = i;
      for (int i = n - 1; i > 0; i--) std::swap(perm[d][i], perm[d][sm(s) % (uint64_t)(i + 1)]);
    }
    for (int i = 0; i < n; i++)
      pts[i] = make_float3(perm[0][i] * 0x1p-20f, perm[1][i] * 0x1p-20f, perm[2][i] * 0x1p-20f);
    for (auto &q : qs)
      q = make_float3((sm(s) >> 40) * 0x1p-24f * n * 0x1p-20f, (sm(s) >> 40) * 0x1p-24f * n * 0x1p-20f,
                      (sm(s) >> 40) * 0x1p-24f * n * 0x1p-20f);
  }
  int bad = 0;

  // 1. host builder vs device thrust builder
  std::vector<float3> host = pts;
  cukd::box_t<float3> hostBox;
  cukd::buildTree_host(host.data(), n, &hostBox);
  float3 *d_pts, *d_q;
  cukd::box_t<float3> *d_wb;
  CHECK(hipMalloc(&d_pts, n * sizeof(float3)));
  CHECK(hipMalloc(&d_q, nq * sizeof(float3)));
  CHECK(hipMalloc(&d_wb, sizeof(*d_wb)));
  CHECK(hipMemcpy(d_q, qs.data(), nq * sizeof(float3), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(d_pts, pts.data(), n * sizeof(float3), hipMemcpyHostToDevice));
  cukd::buildTree_thrust(d_pts, n, d_wb);
  CHECK(hipDeviceSynchronize());
  std::vector<float3> dev(n);
  cukd::box_t<float3> devBox;
  CHECK(hipMemcpy(dev.data(), d_pts, n * sizeof(float3), hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(&devBox, d_wb, sizeof(devBox), hipMemcpyDeviceToHost));
  const int hb = memcmp(host.data(), dev.data(), n * sizeof(float3)) != 0 ||
                 memcmp(&hostBox, &devBox, sizeof(devBox)) != 0;
  printf("buildTree_host equals buildTree_thrust: %s\n", hb ? "no" : "yes");
  bad += hb;

  // 2. compile-time builder selection, seen through the memory resource
  CHECK(hipMemcpy(d_pts, pts.data(), n * sizeof(float3), hipMemcpyHostToDevice));
  CountingResource mr;
  cukd::buildTree(d_pts, n, (cukd::box_t<float3> *)nullptr, 0, mr);
  CHECK(hipDeviceSynchronize());
  CHECK(hipMemcpy(dev.data(), d_pts, n * sizeof(float3), hipMemcpyDeviceToHost));
  const int tb = memcmp(host.data(), dev.data(), n * sizeof(float3)) != 0;
  printf("%s: buildTree made %d allocations (%zu bytes), %d frees; equals host tree: %s\n",
         variant, mr.mallocs, mr.bytes, mr.frees, tb ? "no" : "yes");
  bad += tb;
#if defined(CUKD_BUILDER_INPLACE)
  bad += mr.mallocs != 0 || mr.frees != 0;
#elif defined(CUKD_BUILDER_BITONIC)
  bad += mr.mallocs != 1 || mr.bytes != (size_t)n * 4 || mr.frees != 1;
#else
  size_t ws = 0;
  const cukd_layout L = cukd::layout_of<float3, cukd::default_data_traits<float3>>();
  cukd_build_tree_temp_bytes(n, &L, CUKD_BUILDER_ID_THRUST, &ws);
  bad += mr.mallocs != 1 || mr.bytes != ws || mr.frees != 1;
#endif

  // 3. traversal statistics from a user kernel
  unsigned long long *d_stats, *d_lib;
  CHECK(hipMalloc(&d_stats, sizeof(*d_stats)));
  CHECK(hipMalloc(&d_lib, sizeof(*d_lib)));
  CHECK(hipMemcpyToSymbol(HIP_SYMBOL(cukd::g_traversalStats), &d_stats, sizeof(d_stats)));
  int *d_ids;
  float *d_d2;
  CHECK(hipMalloc(&d_ids, nq * 8 * sizeof(int)));
  CHECK(hipMalloc(&d_d2, nq * 8 * sizeof(float)));
  const int methods[3] = {CUKD_STACK_BASED, CUKD_STACK_FREE, CUKD_CCT};
  for (int m = 0; m < 3; m++) {
    CHECK(hipMemset(d_stats, 0, sizeof(*d_stats)));
    CHECK(hipMemset(d_lib, 0, sizeof(*d_lib)));
    userQueries<<<(nq + 127) / 128, 128>>>(d_pts, n, hostBox, d_q, nq, methods[m]);
    cukd_set_traversal_stats(d_lib);
    cukd::batch::knn<float3>(methods[m], CUKD_LIST_FIXED, 8, d_q, nq, d_pts, n, d_wb, d_ids,
                             d_d2, nullptr, INFINITY, 0, 0);
    cukd_set_traversal_stats(nullptr);
    CHECK(hipDeviceSynchronize());
    unsigned long long user = 0, lib = 0;
    CHECK(hipMemcpy(&user, d_stats, sizeof(user), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(&lib, d_lib, sizeof(lib), hipMemcpyDeviceToHost));
    printf("stats method %d: user kernel %llu visits, library %llu\n", methods[m], user, lib);
    bad += user != lib || user == 0;
  }

  // 4. NaN coordinates are rejected and the array stays as it was
  std::vector<float3> withNan = pts;
  withNan[n / 3].y = NAN;
  CHECK(hipMemcpy(d_pts, withNan.data(), n * sizeof(float3), hipMemcpyHostToDevice));
  bool threw = false;
  try {
    cukd::buildTree(d_pts, n);
  } catch (const std::runtime_error &e) {
    threw = true;
    printf("NaN input: %s\n", e.what());
  }
  CHECK(hipDeviceSynchronize());
  CHECK(hipMemcpy(dev.data(), d_pts, n * sizeof(float3), hipMemcpyDeviceToHost));
  const int nb = !threw || memcmp(withNan.data(), dev.data(), n * sizeof(float3)) != 0;
  bool hthrew = false;
  try {
    cukd::buildTree_host(withNan.data(), n);
  } catch (const std::runtime_error &) {
    hthrew = true;
  }
  printf("NaN rejected by buildTree and buildTree_host, array unchanged: %s\n",
         nb || !hthrew ? "no" : "yes");
  bad += nb || !hthrew;

  CHECK(hipFree(d_pts));
  CHECK(hipFree(d_q));
  CHECK(hipFree(d_wb));
  CHECK(hipFree(d_stats));
  CHECK(hipFree(d_lib));
  CHECK(hipFree(d_ids));
  CHECK(hipFree(d_d2));
  printf(bad ? "FAILED\n" : "OK\n");
  return bad ? 1 : 0;
}
