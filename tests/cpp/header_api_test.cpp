// Exercises the drop-in C++ header API the way the reference's samples
// and tests use it: typed data_traits (payload + explicit dim, uint8 dim),
// device-side fcp/knn called from a user kernel, the batched calls, and the
// spatial tree.  Device results must equal the batched library results.
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <cmath>
#include <vector>

#include "cukd/kdtree.h"

#define CHECK(x)                                                           \
  do {                                                                     \
    hipError_t e = (x);                                                    \
    if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } \
  } while (0)

// testing/testBuildersSameResult.cu-style payload type with a split dim
struct PointPlusPayload {
  float3 pos;
  int payload;
  int splitDim;
};
struct PointPlusPayload_traits : public cukd::default_data_traits<float3> {
  using data_t = PointPlusPayload;
  using point_t = float3;
  enum { has_explicit_dim = true };
  static inline __both__ const float3 &get_point(const data_t &d) { return d.pos; }
  static inline __both__ float get_coord(const data_t &d, int k) { return cukd::get_coord(d.pos, k); }
  static inline __both__ int get_dim(const data_t &d) { return d.splitDim; }
  static inline __both__ void set_dim(data_t &d, int k) { d.splitDim = k; }
};

// testing/testBuilderSimpleInput.cu-style Photon with a uint8_t split dim
struct Photon {
  float3 position;
  float3 power;
  uint8_t splitDim;
  uint8_t pad[3];
};
struct Photon_traits : public cukd::default_data_traits<float3> {
  using data_t = Photon;
  using point_t = float3;
  enum { has_explicit_dim = true };
  static inline __both__ const float3 &get_point(const Photon &p) { return p.position; }
  static inline __both__ float get_coord(const Photon &p, int k) { return cukd::get_coord(p.position, k); }
  static inline __both__ int get_dim(const Photon &p) { return p.splitDim; }
  static inline __both__ void set_dim(Photon &p, int k) { p.splitDim = (uint8_t)k; }
};

__global__ void userQueries(const float3 *tree, int n, cukd::box_t<float3> wb, const float3 *q,
                            int nq, int *fcpOut, int *knnOut)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  fcpOut[3 * i + 0] = cukd::stackBased::fcp(q[i], tree, n);
  fcpOut[3 * i + 1] = cukd::stackFree::fcp(q[i], tree, n);
  fcpOut[3 * i + 2] = cukd::cct::fcp(q[i], wb, tree, n);
  cukd::FixedCandidateList<8> fixed(INFINITY);
  cukd::stackBased::knn(fixed, q[i], tree, n);
  cukd::HeapCandidateList<8> heap(INFINITY);
  cukd::cct::knn(heap, q[i], wb, tree, n);
  for (int j = 0; j < 8; j++) {
    knnOut[16 * i + j] = fixed.get_pointID(j);
    knnOut[16 * i + 8 + j] = heap.get_pointID(j);
  }
}

__global__ void photonQueries(const Photon *tree, int n, const float3 *q, int nq, int *out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  out[i] = cukd::stackBased::fcp<Photon, Photon_traits>(q[i], tree, n);
}

__global__ void spatialQueries(cukd::SpatialKDTree<float3> tree, const float3 *q, int nq, int *out)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nq) return;
  out[2 * i] = cukd::stackBased::fcp(tree, q[i]);
  out[2 * i + 1] = cukd::cct::fcp(tree, q[i]);
}

static uint64_t sm(uint64_t &x)
{
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static float u01(uint64_t &x) { return (float)((sm(x) >> 40) * (1.0 / (1 << 24))); }

int main()
{
  const int n = 100000, nq = 4096;
  uint64_t s = 1;
  std::vector<float3> pts(n), qs(nq);
  for (auto &p : pts) p = make_float3(u01(s), u01(s), u01(s));
  for (auto &p : qs) p = make_float3(u01(s), u01(s), u01(s));
  float3 *d_pts, *d_q;
  cukd::box_t<float3> *d_wb;
  int *d_a, *d_b, *d_c;
  CHECK(hipMalloc(&d_pts, n * sizeof(float3)));
  CHECK(hipMalloc(&d_q, nq * sizeof(float3)));
  CHECK(hipMalloc(&d_wb, sizeof(*d_wb)));
  CHECK(hipMalloc(&d_a, nq * 16 * sizeof(int)));
  CHECK(hipMalloc(&d_b, nq * 16 * sizeof(int)));
  CHECK(hipMalloc(&d_c, nq * 16 * sizeof(int)));
  CHECK(hipMemcpy(d_pts, pts.data(), n * sizeof(float3), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(d_q, qs.data(), nq * sizeof(float3), hipMemcpyHostToDevice));

  // 1. float3 default traits
  cukd::buildTree(d_pts, n, d_wb);
  CHECK(hipDeviceSynchronize());
  cukd::box_t<float3> wb;
  CHECK(hipMemcpy(&wb, d_wb, sizeof(wb), hipMemcpyDeviceToHost));
  userQueries<<<(nq + 127) / 128, 128>>>(d_pts, n, wb, d_q, nq, d_a, d_b);
  CHECK(hipDeviceSynchronize());
  std::vector<int> fcpDev(3 * nq), knnDev(16 * nq), ref(16 * nq);
  CHECK(hipMemcpy(fcpDev.data(), d_a, 3 * nq * sizeof(int), hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(knnDev.data(), d_b, 16 * nq * sizeof(int), hipMemcpyDeviceToHost));
  int bad = 0;
  const int methods[3] = {CUKD_STACK_BASED, CUKD_STACK_FREE, CUKD_CCT};
  for (int m = 0; m < 3; m++) {
    cukd::batch::fcp<float3>(methods[m], d_q, nq, d_pts, n, d_wb, d_c);
    CHECK(hipDeviceSynchronize());
    CHECK(hipMemcpy(ref.data(), d_c, nq * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < nq; i++) bad += ref[i] != fcpDev[3 * i + m];
  }
  float *d_d2;
  CHECK(hipMalloc(&d_d2, nq * 8 * sizeof(float)));
  cukd::batch::knn<float3>(CUKD_STACK_BASED, CUKD_LIST_FIXED, 8, d_q, nq, d_pts, n, d_wb, d_c, d_d2);
  CHECK(hipDeviceSynchronize());
  CHECK(hipMemcpy(ref.data(), d_c, 8 * nq * sizeof(int), hipMemcpyDeviceToHost));
  for (int i = 0; i < nq; i++)
    for (int j = 0; j < 8; j++) bad += ref[8 * i + j] != knnDev[16 * i + j];
  cukd::batch::knn<float3>(CUKD_CCT, CUKD_LIST_HEAP, 8, d_q, nq, d_pts, n, d_wb, d_c, d_d2);
  CHECK(hipDeviceSynchronize());
  CHECK(hipMemcpy(ref.data(), d_c, 8 * nq * sizeof(int), hipMemcpyDeviceToHost));
  for (int i = 0; i < nq; i++)
    for (int j = 0; j < 8; j++) bad += ref[8 * i + j] != knnDev[16 * i + 8 + j];
  printf("float3: device-vs-batch mismatches %d\n", bad);

  // 2. payload + explicit dims through the traits probe
  cukd_layout L = cukd::layout_of<PointPlusPayload, PointPlusPayload_traits>();
  if (L.stride != 20 || L.point_offset != 0 || L.dim_offset != 16 || L.dim_bytes != 4) {
    printf("bad PointPlusPayload layout\n");
    bad++;
  }
  std::vector<PointPlusPayload> pp(n);
  for (int i = 0; i < n; i++) pp[i] = {pts[i], i, -1};
  PointPlusPayload *d_pp;
  CHECK(hipMalloc(&d_pp, n * sizeof(PointPlusPayload)));
  CHECK(hipMemcpy(d_pp, pp.data(), n * sizeof(PointPlusPayload), hipMemcpyHostToDevice));
  bool threw = false;
  try {
    cukd::buildTree<PointPlusPayload, PointPlusPayload_traits>(d_pp, n);
  } catch (const std::runtime_error &) {
    threw = true;
  }
  if (!threw) { printf("explicit dims without bounds did not throw\n"); bad++; }
  cukd::buildTree<PointPlusPayload, PointPlusPayload_traits>(d_pp, n, d_wb);
  CHECK(hipDeviceSynchronize());
  CHECK(hipMemcpy(pp.data(), d_pp, n * sizeof(PointPlusPayload), hipMemcpyDeviceToHost));
  std::vector<int> seen(n, 0);
  for (int i = 0; i < n; i++) {
    if (pp[i].payload >= 0 && pp[i].payload < n) seen[pp[i].payload]++;
    if (pp[i].splitDim < 0 || pp[i].splitDim > 2) bad++;
  }
  for (int i = 0; i < n; i++) bad += seen[i] != 1;
  printf("PointPlusPayload: dims set, payload permutation ok (bad %d)\n", bad);

  // 3. Photon with uint8_t dim
  L = cukd::layout_of<Photon, Photon_traits>();
  if (L.stride != 28 || L.dim_offset != 24 || L.dim_bytes != 1) { printf("bad Photon layout\n"); bad++; }
  std::vector<Photon> ph(n);
  for (int i = 0; i < n; i++) { ph[i].position = pts[i]; ph[i].power = make_float3(i, 0, 0); ph[i].splitDim = 7; }
  Photon *d_ph;
  CHECK(hipMalloc(&d_ph, n * sizeof(Photon)));
  CHECK(hipMemcpy(d_ph, ph.data(), n * sizeof(Photon), hipMemcpyHostToDevice));
  cukd::buildTree<Photon, Photon_traits>(d_ph, n, d_wb);
  photonQueries<<<(nq + 127) / 128, 128>>>(d_ph, n, d_q, nq, d_a);
  CHECK(hipDeviceSynchronize());
  cukd::batch::fcp<Photon, Photon_traits>(CUKD_STACK_BASED, d_q, nq, d_ph, n, d_wb, d_c);
  CHECK(hipDeviceSynchronize());
  CHECK(hipMemcpy(fcpDev.data(), d_a, nq * sizeof(int), hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(ref.data(), d_c, nq * sizeof(int), hipMemcpyDeviceToHost));
  for (int i = 0; i < nq; i++) bad += ref[i] != fcpDev[i];
  printf("Photon: device-vs-batch (bad %d)\n", bad);

  // 4. spatial tree
  CHECK(hipMemcpy(d_pts, pts.data(), n * sizeof(float3), hipMemcpyHostToDevice));
  cukd::SpatialKDTree<float3> st;
  cukd::buildTree(st, d_pts, n);
  spatialQueries<<<(nq + 127) / 128, 128>>>(st, d_q, nq, d_a);
  CHECK(hipDeviceSynchronize());
  CHECK(hipMemcpy(fcpDev.data(), d_a, 2 * nq * sizeof(int), hipMemcpyDeviceToHost));
  for (int i = 0; i < nq; i++) {
    // both traversals find a point at the brute-force minimum distance
    float best = INFINITY;
    for (int j = 0; j < n; j += 1) {
      float dx = pts[j].x - qs[i].x, dy = pts[j].y - qs[i].y, dz = pts[j].z - qs[i].z;
      float d = dx * dx + dy * dy + dz * dz;
      best = d < best ? d : best;
    }
    for (int m = 0; m < 2; m++) {
      const int id = fcpDev[2 * i + m];
      float dx = pts[id].x - qs[i].x, dy = pts[id].y - qs[i].y, dz = pts[id].z - qs[i].z;
      bad += (dx * dx + dy * dy + dz * dz) != best;
    }
    if (i == 255) break;
  }
  cukd::free(st);
  printf("spatial: nodes %d (bad %d)\n", st.numNodes, bad);

  // 5. forest: three trees built together equal three buildTree calls, and
  //    forest kNN equals per-tree batch kNN
  {
    const std::vector<int64_t> off = {0, 1000, 1000, n};  // one tree empty
    float3 *d_f;
    CHECK(hipMalloc(&d_f, n * sizeof(float3)));
    CHECK(hipMemcpy(d_f, pts.data(), n * sizeof(float3), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(d_pts, pts.data(), n * sizeof(float3), hipMemcpyHostToDevice));
    cukd::Forest<float3> forest(off);
    forest.build(d_f);
    cukd::buildTree(d_pts, 1000);
    cukd::buildTree(d_pts + 1000, n - 1000);
    CHECK(hipDeviceSynchronize());
    std::vector<float3> a(n), b(n);
    CHECK(hipMemcpy(a.data(), d_f, n * sizeof(float3), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(b.data(), d_pts, n * sizeof(float3), hipMemcpyDeviceToHost));
    int fb = memcmp(a.data(), b.data(), n * sizeof(float3)) != 0;
    // queries: first half on tree 2, second half on tree 0
    const int segOff[3] = {0, nq / 2, nq};
    const int segTree[2] = {2, 0};
    int *d_so, *d_st;
    CHECK(hipMalloc(&d_so, sizeof(segOff)));
    CHECK(hipMalloc(&d_st, sizeof(segTree)));
    CHECK(hipMemcpy(d_so, segOff, sizeof(segOff), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(d_st, segTree, sizeof(segTree), hipMemcpyHostToDevice));
    forest.knn(CUKD_STACK_BASED, CUKD_LIST_FIXED, 4, d_f, nullptr, d_q, nq, 2, d_so, d_st, d_a,
               d_d2);
    cukd::batch::knn<float3>(CUKD_STACK_BASED, CUKD_LIST_FIXED, 4, d_q, nq / 2, d_pts + 1000,
                             n - 1000, nullptr, d_c, d_d2 + 4 * nq);
    cukd::batch::knn<float3>(CUKD_STACK_BASED, CUKD_LIST_FIXED, 4, d_q + nq / 2, nq - nq / 2,
                             d_pts, 1000, nullptr, d_c + 2 * nq, d_d2 + 4 * nq);
    CHECK(hipDeviceSynchronize());
    std::vector<int> x(4 * nq), y(4 * nq);
    CHECK(hipMemcpy(x.data(), d_a, 4 * nq * sizeof(int), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(y.data(), d_c, 4 * nq * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < 4 * nq; i++) fb += x[i] != y[i];
    bad += fb;
    printf("forest: build and kNN equal per-tree calls (bad %d)\n", fb);

    // 6. two accumulate rounds over the two halves of the points equal one
    //    heap-list kNN over all of them (as sets; ids made global)
    uint64_t *d_lists;
    CHECK(hipMalloc(&d_lists, (size_t)nq * 8 * sizeof(uint64_t)));
    cukd::knnAccumulate<float3>(CUKD_STACK_FREE, 8, d_f + 1000, n - 1000, nullptr, d_q, nq,
                                d_lists, true, INFINITY, 1000);
    cukd::knnAccumulate<float3>(CUKD_STACK_FREE, 8, d_f, 1000, nullptr, d_q, nq, d_lists, false);
    float *d_r;
    CHECK(hipMalloc(&d_r, nq * sizeof(float)));
    cukd::knnExtract(d_lists, nq, 8, d_a, d_d2, d_r);
    CHECK(hipDeviceSynchronize());
    std::vector<float> d2a(8 * nq), rad(nq);
    CHECK(hipMemcpy(d2a.data(), d_d2, 8 * nq * sizeof(float), hipMemcpyDeviceToHost));
    CHECK(hipMemcpy(rad.data(), d_r, nq * sizeof(float), hipMemcpyDeviceToHost));
    int mb = 0;
    for (int i = 0; i < 64; i++) {
      // brute force over the original points: the 8 smallest dist2
      std::vector<float> all(n);
      for (int j = 0; j < n; j++) {
        float dx = pts[j].x - qs[i].x, dy = pts[j].y - qs[i].y, dz = pts[j].z - qs[i].z;
        all[j] = dx * dx + dy * dy + dz * dz;
      }
      std::partial_sort(all.begin(), all.begin() + 8, all.end());
      for (int j = 0; j < 8; j++) mb += all[j] != d2a[8 * i + j];
      mb += rad[i] != sqrtf(all[7]);
    }
    bad += mb;
    printf("multi-round: two rounds equal brute force (bad %d)\n", mb);
    CHECK(hipFree(d_lists));
    CHECK(hipFree(d_r));
    CHECK(hipFree(d_f));
    CHECK(hipFree(d_so));
    CHECK(hipFree(d_st));
  }
  printf(bad ? "FAILED\n" : "OK\n");
  return bad ? 1 : 0;
}
