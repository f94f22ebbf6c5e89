// CPU kernel-emulator test (test infrastructure): the library's unmodified
// kernels, compiled with g++ over include/hip/ of this folder and built with
// AddressSanitizer and UndefinedBehaviorSanitizer, run through the C ABI on
// small inputs and are compared with the oracle (oracle/cukd_oracle.c).
// An out-of-bounds access, a use of freed memory or undefined behaviour in
// any kernel or host path aborts the run with the sanitizer's report; a
// result that differs from the oracle fails it.  This is where faults are
// found: on the CPU, before any GPU run.
//
//   emu_test [case-name-substring]
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "cukd_amd.h"
#include "../../oracle/cukd_oracle.h"

// query kernels are compiled for dims 2, 3 and 4 only
namespace cukd_amd {
struct QueryArgs;
#define CUKD_EMU_NO_DIM(d)                                                          \
  hipError_t launch_query_##d(int, const QueryArgs &, hipStream_t) { return hipErrorInvalidValue; }
CUKD_EMU_NO_DIM(1)
CUKD_EMU_NO_DIM(5)
CUKD_EMU_NO_DIM(6)
CUKD_EMU_NO_DIM(7)
CUKD_EMU_NO_DIM(8)
}  // namespace cukd_amd

namespace {

uint64_t sm(uint64_t &x)
{
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// n points of D coordinates packed into stride-byte elements
std::vector<uint32_t> points(int n, const cukd_layout &L, uint64_t seed, const char *dist)
{
  const int W = L.stride / 4, D = L.num_dims, po = L.point_offset / 4;
  std::vector<uint32_t> e((size_t)n * W);
  for (size_t i = 0; i < e.size(); i++) e[i] = (uint32_t)sm(seed);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < D; d++) {
      const uint64_t r = sm(seed);
      float f;
      if (!strcmp(dist, "ties")) f = (float)(r % 7);
      else if (!strcmp(dist, "dups")) f = 0.5f;
      else if (!strcmp(dist, "twovals")) f = (float)(r & 1);
      else if (!strcmp(dist, "coarse")) f = (float)(r % 4096) / 4096.f;
      else if (!strcmp(dist, "fine")) f = (float)(r % 16384) / 16384.f;  // short tie runs
      else if (!strcmp(dist, "grid17")) f = (float)(r % 131072) / 131072.f;  // sparser ties
      else f = (float)(r >> 40) * (1.f / 16777216.f);
      uint32_t b;
      if (L.scalar_type == CUKD_SCALAR_F32) memcpy(&b, &f, 4);
      else b = (uint32_t)(int32_t)(f * 1000.f);
      e[(size_t)i * W + po + d] = b;
    }
  return e;
}

std::vector<float> queries(int nq, int D, uint64_t seed)
{
  std::vector<float> q((size_t)nq * D);
  for (auto &v : q) v = (float)(sm(seed) >> 40) * (1.f / 16777216.f) * 1.1f - 0.05f;
  return q;
}

oc_layout ocl(const cukd_layout &L)
{
  return oc_layout{L.num_dims, L.scalar_type, L.stride, L.point_offset, L.dim_offset, L.dim_bytes};
}

int g_fail = 0;

void report(const std::string &name, bool ok)
{
  printf("%-58s %s\n", name.c_str(), ok ? "ok" : "FAILED");
  fflush(stdout);
  if (!ok) g_fail++;
}

bool build_case(const char *dist, int n, const cukd_layout &L, int builder, uint64_t seed)
{
  std::vector<uint32_t> e = points(n, L, seed, dist);
  std::vector<uint32_t> ref = e;
  std::vector<uint32_t> rbox(2 * L.num_dims), gbox(2 * L.num_dims);
  const oc_layout O = ocl(L);
  if (oc_build_host(ref.data(), n, &O, rbox.data()) != 0) return false;
  void *d = nullptr;
  hipMalloc(&d, e.size() * 4 + 4);
  memcpy(d, e.data(), e.size() * 4);
  void *db = nullptr;
  hipMalloc(&db, gbox.size() * 4);
  const int rc = cukd_build_tree(d, n, &L, builder, db, nullptr, nullptr);
  memcpy(gbox.data(), db, gbox.size() * 4);
  const bool ok = rc == CUKD_OK && !memcmp(d, ref.data(), e.size() * 4) &&
                  (n < 1 || !memcmp(gbox.data(), rbox.data(), gbox.size() * 4));
  if (rc != CUKD_OK) printf("  build failed: %s\n", cukd_last_error());
  hipFree(d);
  hipFree(db);
  return ok;
}

bool query_case(int n, int nq, int D, int method, int k, int flags, uint64_t seed,
                int list = CUKD_LIST_FIXED)
{
  cukd_layout L{D, CUKD_SCALAR_F32, 4 * D, 0, -1, 4};
  std::vector<uint32_t> e = points(n, L, seed, "coarse");
  const oc_layout O = ocl(L);
  std::vector<float> box(2 * D);
  oc_build_host(e.data(), n, &O, box.data());
  std::vector<float> q = queries(nq, D, seed + 1);
  const int kk = k > 0 ? k : 1;
  std::vector<int32_t> rid((size_t)nq * kk), gid((size_t)nq * kk);
  std::vector<float> rd2((size_t)nq * kk), gd2((size_t)nq * kk), rr(nq), gr(nq);
  void *dt, *dq, *db, *di, *dd, *dr;
  hipMalloc(&dt, e.size() * 4 + 4);
  hipMalloc(&dq, q.size() * 4 + 4);
  hipMalloc(&db, box.size() * 4);
  hipMalloc(&di, rid.size() * 4 + 4);
  hipMalloc(&dd, rd2.size() * 4 + 4);
  hipMalloc(&dr, rr.size() * 4 + 4);
  memcpy(dt, e.data(), e.size() * 4);
  memcpy(dq, q.data(), q.size() * 4);
  memcpy(db, box.data(), box.size() * 4);
  int rc;
  if (k == 0) {
    oc_fcp(method, e.data(), n, &O, box.data(), q.data(), nq, INFINITY, 0.f, rid.data(),
           rd2.data(), nullptr, 1);
    rc = cukd_fcp(method, dt, n, &L, db, (const float *)dq, nq, INFINITY, 0.f, (int32_t *)di,
                  (float *)dd, flags, nullptr, nullptr);
  } else {
    oc_knn(method, list, k, e.data(), n, &O, box.data(), q.data(), nq, INFINITY,
           rid.data(), rd2.data(), rr.data(), nullptr, 1);
    rc = (flags & 64) ? cukd_knn_noalloc(method, list, k, dt, n, &L, db,
                                          (const float *)dq, nq, INFINITY, (int32_t *)di,
                                          (float *)dd, (float *)dr, flags & ~64, nullptr)
                      : cukd_knn(method, list, k, dt, n, &L, db, (const float *)dq,
                                 nq, INFINITY, (int32_t *)di, (float *)dd, (float *)dr, flags,
                                 nullptr, nullptr);
    memcpy(gr.data(), dr, rr.size() * 4);
  }
  memcpy(gid.data(), di, gid.size() * 4);
  memcpy(gd2.data(), dd, gd2.size() * 4);
  bool ok = rc == CUKD_OK && gid == rid && !memcmp(gd2.data(), rd2.data(), gd2.size() * 4);
  if (k > 0) ok = ok && !memcmp(gr.data(), rr.data(), rr.size() * 4);
  if (rc != CUKD_OK) printf("  query failed: %s\n", cukd_last_error());
  for (void *p : {dt, dq, db, di, dd, dr}) hipFree(p);
  return ok;
}

bool forest_case(const std::vector<int64_t> &off, uint64_t seed, int per = 40)
{
  cukd_layout L{3, CUKD_SCALAR_F32, 12, 0, -1, 4};
  const int T = (int)off.size() - 1, N = (int)off.back();
  std::vector<uint32_t> e = points(N, L, seed, "coarse");
  std::vector<uint32_t> ref = e;
  const oc_layout O = ocl(L);
  for (int t = 0; t < T; t++)
    if (off[t + 1] > off[t]) oc_build_host(ref.data() + 3 * off[t], off[t + 1] - off[t], &O, nullptr);
  void *d;
  hipMalloc(&d, e.size() * 4 + 4);
  memcpy(d, e.data(), e.size() * 4);
  cukd_forest *f = nullptr;
  int rc = cukd_forest_create(&f, off.data(), T, &L, nullptr, nullptr);
  if (rc == CUKD_OK) rc = cukd_forest_build(f, d, nullptr, nullptr);
  bool ok = rc == CUKD_OK && !memcmp(d, ref.data(), e.size() * 4);
  // kNN-4 over every tree, per queries each, against the oracle
  const int nq = per * T, k = 4;
  std::vector<float> q = queries(nq, 3, seed + 2);
  std::vector<int32_t> so(T + 1), st(T);
  for (int t = 0; t <= T; t++) so[t] = per * t;
  for (int t = 0; t < T; t++) st[t] = t;
  void *dq, *dso, *dst, *di, *dd;
  hipMalloc(&dq, q.size() * 4);
  hipMalloc(&dso, so.size() * 4);
  hipMalloc(&dst, st.size() * 4);
  hipMalloc(&di, (size_t)nq * k * 4);
  hipMalloc(&dd, (size_t)nq * k * 4);
  memcpy(dq, q.data(), q.size() * 4);
  memcpy(dso, so.data(), so.size() * 4);
  memcpy(dst, st.data(), st.size() * 4);
  if (ok) {
    rc = cukd_forest_knn(CUKD_STACK_BASED, CUKD_LIST_FIXED, k, f, d, nullptr, (const float *)dq,
                         nq, T, (const int32_t *)dso, (const int32_t *)dst, INFINITY,
                         (int32_t *)di, (float *)dd, nullptr, CUKD_QUERY_SORT, nullptr);
    ok = rc == CUKD_OK;
    for (int t = 0; t < T && ok; t++) {
      std::vector<int32_t> rid(per * k);
      std::vector<float> rd2(per * k);
      oc_knn(OC_STACK_BASED, OC_LIST_FIXED, k, ref.data() + 3 * off[t], off[t + 1] - off[t], &O,
             nullptr, q.data() + 3 * per * t, per, INFINITY, rid.data(), rd2.data(), nullptr,
             nullptr, 1);
      ok = !memcmp(rid.data(), (int32_t *)di + per * k * t, rid.size() * 4) &&
           !memcmp(rd2.data(), (float *)dd + per * k * t, rd2.size() * 4);
    }
  }
  if (rc != CUKD_OK) printf("  forest failed: %s\n", cukd_last_error());
  if (f) cukd_forest_destroy(f, nullptr);
  for (void *p : {d, dq, dso, dst, di, dd}) hipFree(p);
  return ok;
}

// two shards into one set of persistent heap lists, against the oracle's
// lists word for word, then the extraction
bool accumulate_case(int n, int nq, int k, int method, uint64_t seed)
{
  cukd_layout L{3, CUKD_SCALAR_F32, 12, 0, -1, 4};
  const oc_layout O = ocl(L);
  std::vector<uint32_t> s0 = points(n, L, seed, "coarse"), s1 = points(n / 2 + 1, L, seed + 7, "coarse");
  std::vector<float> b0(6), b1(6);
  oc_build_host(s0.data(), n, &O, b0.data());
  oc_build_host(s1.data(), n / 2 + 1, &O, b1.data());
  std::vector<float> q = queries(nq, 3, seed + 1);
  std::vector<uint64_t> rl((size_t)nq * k), gl((size_t)nq * k);
  oc_knn_accumulate(method, k, s0.data(), n, &O, b0.data(), q.data(), nq, rl.data(), 1, 0.3f, 0,
                    nullptr, 1);
  oc_knn_accumulate(method, k, s1.data(), n / 2 + 1, &O, b1.data(), q.data(), nq, rl.data(), 0,
                    0.3f, n, nullptr, 1);
  void *d0, *d1, *db0, *db1, *dq, *dl, *di, *dd, *dr;
  hipMalloc(&d0, s0.size() * 4);
  hipMalloc(&d1, s1.size() * 4);
  hipMalloc(&db0, 24);
  hipMalloc(&db1, 24);
  hipMalloc(&dq, q.size() * 4);
  hipMalloc(&dl, rl.size() * 8);
  hipMalloc(&di, (size_t)nq * k * 4);
  hipMalloc(&dd, (size_t)nq * k * 4);
  hipMalloc(&dr, (size_t)nq * 4);
  memcpy(d0, s0.data(), s0.size() * 4);
  memcpy(d1, s1.data(), s1.size() * 4);
  memcpy(db0, b0.data(), 24);
  memcpy(db1, b1.data(), 24);
  memcpy(dq, q.data(), q.size() * 4);
  int rc = cukd_knn_accumulate(method, k, d0, n, &L, db0, (const float *)dq, nq, (uint64_t *)dl, 1,
                               0.3f, 0, 0, nullptr, nullptr);
  if (rc == CUKD_OK)
    rc = cukd_knn_accumulate(method, k, d1, n / 2 + 1, &L, db1, (const float *)dq, nq,
                             (uint64_t *)dl, 0, 0.3f, n, 0, nullptr, nullptr);
  memcpy(gl.data(), dl, gl.size() * 8);
  bool ok = rc == CUKD_OK && gl == rl;
  if (ok) {
    std::vector<int32_t> rid((size_t)nq * k);
    std::vector<float> rd2((size_t)nq * k), rr(nq);
    oc_knn_extract(rl.data(), nq, k, rid.data(), rd2.data(), rr.data());
    rc = cukd_knn_extract((const uint64_t *)dl, nq, k, (int32_t *)di, (float *)dd, (float *)dr,
                          nullptr);
    ok = rc == CUKD_OK && !memcmp(di, rid.data(), rid.size() * 4) &&
         !memcmp(dd, rd2.data(), rd2.size() * 4) && !memcmp(dr, rr.data(), rr.size() * 4);
  }
  if (rc != CUKD_OK) printf("  accumulate failed: %s\n", cukd_last_error());
  for (void *p : {d0, d1, db0, db1, dq, dl, di, dd, dr}) hipFree(p);
  return ok;
}

bool spatial_case(int n, int D, int scalar, uint64_t seed, int k = 8, int list = CUKD_LIST_FIXED)
{
  cukd_layout L{D, scalar, 4 * D, 0, -1, 4};
  std::vector<uint32_t> e = points(n, L, seed, "coarse");
  const oc_layout O = ocl(L);
  const int64_t nn = oc_spatial_build(e.data(), n, &O, 0, nullptr, nullptr, nullptr);
  std::vector<oc_spatial_node> rn(nn);
  std::vector<uint32_t> rp(n > 0 ? n : 1), rb(2 * D);
  oc_spatial_build(e.data(), n, &O, 0, rn.data(), rp.data(), (float *)rb.data());
  void *d;
  hipMalloc(&d, e.size() * 4 + 4);
  memcpy(d, e.data(), e.size() * 4);
  cukd_spatial_tree t;
  int rc = cukd_spatial_build(&t, d, n, &L, 0, nullptr, nullptr);
  bool ok = rc == CUKD_OK && t.num_nodes == nn &&
            !memcmp(t.nodes, rn.data(), nn * sizeof(oc_spatial_node)) &&
            !memcmp(t.prim_ids, rp.data(), (size_t)n * 4) &&
            !memcmp(t.ibounds, rb.data(), rb.size() * 4);
  if (ok && scalar == CUKD_SCALAR_F32) {
    const int nq = 500;
    std::vector<float> q = queries(nq, D, seed + 3);
    std::vector<int32_t> rid(nq * k), gid(nq * k);
    std::vector<float> rd2(nq * k), gd2(nq * k);
    oc_spatial_knn(OC_STACK_BASED, list, k, rn.data(), rp.data(), (float *)rb.data(),
                   e.data(), &O, q.data(), nq, 0.2f, rid.data(), rd2.data(), nullptr, 1);
    void *dq, *di, *dd;
    hipMalloc(&dq, q.size() * 4);
    hipMalloc(&di, gid.size() * 4);
    hipMalloc(&dd, gd2.size() * 4);
    memcpy(dq, q.data(), q.size() * 4);
    rc = cukd_spatial_knn(CUKD_STACK_BASED, list, k, &t, &L, (const float *)dq, nq,
                          0.2f, (int32_t *)di, (float *)dd, nullptr, nullptr);
    ok = rc == CUKD_OK && !memcmp(di, rid.data(), gid.size() * 4) &&
         !memcmp(dd, rd2.data(), gd2.size() * 4);
    for (void *p : {dq, di, dd}) hipFree(p);
  }
  if (rc != CUKD_OK) printf("  spatial failed: %s\n", cukd_last_error());
  cukd_spatial_free(&t, nullptr, nullptr);
  hipFree(d);
  return ok;
}

}  // namespace

int main(int argc, char **argv)
{
  const std::string only = argc > 1 ? argv[1] : "";
  struct Case {
    std::string name;
    std::function<bool()> run;
  };
  const cukd_layout f3{3, CUKD_SCALAR_F32, 12, 0, -1, 4};
  const cukd_layout f2{2, CUKD_SCALAR_F32, 8, 0, -1, 4};
  const cukd_layout i4{4, CUKD_SCALAR_I32, 16, 0, -1, 4};
  const cukd_layout ppp{3, CUKD_SCALAR_F32, 20, 0, 16, 4};   // PointPlusPayload + dim
  const cukd_layout photon{3, CUKD_SCALAR_F32, 16, 0, 12, 1};
  std::vector<Case> cases;
  for (int n : {1, 2, 5, 100, 1000, 9000, 40000})
    cases.push_back({"build default float3 uniform n=" + std::to_string(n),
                     [=] { return build_case("uniform", n, f3, CUKD_BUILDER_ID_DEFAULT, 11 + n); }});
  for (int n : {3000, 20000})
    cases.push_back({"build default float3 ties n=" + std::to_string(n),
                     [=] { return build_case("ties", n, f3, CUKD_BUILDER_ID_DEFAULT, 13 + n); }});
  // tie groups that span whole subtrees: every key word tied down to the
  // index (dups), or few distinct points (twovals)
  cases.push_back({"build default float3 dups n=20000",
                   [=] { return build_case("dups", 20000, f3, CUKD_BUILDER_ID_DEFAULT, 14); }});
  cases.push_back({"build default float3 twovals n=20000",
                   [=] { return build_case("twovals", 20000, f3, CUKD_BUILDER_ID_DEFAULT, 15); }});
  cases.push_back({"build explicit-dim payload dups n=9000",
                   [=] { return build_case("dups", 9000, ppp, CUKD_BUILDER_ID_DEFAULT, 16); }});
  cases.push_back({"build default float2 coarse n=30000",
                   [=] { return build_case("coarse", 30000, f2, CUKD_BUILDER_ID_DEFAULT, 17); }});
  cases.push_back({"build default int4 ties n=12000",
                   [=] { return build_case("ties", 12000, i4, CUKD_BUILDER_ID_DEFAULT, 19); }});
  for (int n : {1000, 9000})
    cases.push_back({"build explicit-dim payload n=" + std::to_string(n),
                     [=] { return build_case("uniform", n, ppp, CUKD_BUILDER_ID_DEFAULT, 23 + n); }});
  cases.push_back({"build explicit-dim photon ties n=5000",
                   [=] { return build_case("ties", 5000, photon, CUKD_BUILDER_ID_DEFAULT, 29); }});
  // short coordinate tie runs: the rank finish's chained modes repair them
  // level by level (explicit dims; round-robin subtrees whose roots lie
  // above level D-1, which the bigcap build gives n < 16384)
  // a few percent of the elements tied: below the hand-over density, so
  // the rank finish repairs the runs itself
  cases.push_back({"build explicit-dim payload grid17 n=20000",
                   [=] { return build_case("grid17", 20000, ppp, CUKD_BUILDER_ID_DEFAULT, 97); }});
  cases.push_back({"build explicit-dim photon grid17 n=7000",
                   [=] { return build_case("grid17", 7000, photon, CUKD_BUILDER_ID_DEFAULT, 98); }});
  cases.push_back({"build default float3 grid17 n=12000",
                   [=] { return build_case("grid17", 12000, f3, CUKD_BUILDER_ID_DEFAULT, 99); }});
  cases.push_back({"build explicit-dim payload fine n=20000",
                   [=] { return build_case("fine", 20000, ppp, CUKD_BUILDER_ID_DEFAULT, 91); }});
  cases.push_back({"build explicit-dim float4 payload fine n=9000", [=] {
                     const cukd_layout ppp4{4, CUKD_SCALAR_F32, 24, 0, 20, 4};
                     return build_case("fine", 9000, ppp4, CUKD_BUILDER_ID_DEFAULT, 96);
                   }});
  cases.push_back({"build explicit-dim photon fine n=7000",
                   [=] { return build_case("fine", 7000, photon, CUKD_BUILDER_ID_DEFAULT, 92); }});
  cases.push_back({"build default float3 fine n=12000",
                   [=] { return build_case("fine", 12000, f3, CUKD_BUILDER_ID_DEFAULT, 93); }});
  cases.push_back({"build default float2 fine n=3000",
                   [=] { return build_case("fine", 3000, f2, CUKD_BUILDER_ID_DEFAULT, 94); }});
  cases.push_back({"build default int4 fine n=12000",
                   [=] { return build_case("fine", 12000, i4, CUKD_BUILDER_ID_DEFAULT, 95); }});
  cases.push_back({"build bitonic float3 ties n=5000",
                   [=] { return build_case("ties", 5000, f3, CUKD_BUILDER_ID_BITONIC, 31); }});
  cases.push_back({"build in-place float3 uniform n=5000",
                   [=] { return build_case("uniform", 5000, f3, CUKD_BUILDER_ID_INPLACE, 37); }});
  for (int m : {CUKD_STACK_BASED, CUKD_STACK_FREE, CUKD_CCT})
    for (int k : {0, 1, 8, 20})
      cases.push_back({"query method=" + std::to_string(m) + " k=" + std::to_string(k),
                       [=] { return query_case(6000, 3000, 3, m, k, 0, 41 + m * 10 + k); }});
  // run-time-k lists (entries in LDS columns): fixed k = 20 / 64, heap k = 5 / 64
  for (int m : {CUKD_STACK_BASED, CUKD_STACK_FREE, CUKD_CCT}) {
    cases.push_back({"query method=" + std::to_string(m) + " fixed k=64", [=] {
                       return query_case(3000, 700, 3, m, 64, 0, 111 + m);
                     }});
    for (int k : {5, 64})
      cases.push_back({"query method=" + std::to_string(m) + " heap k=" + std::to_string(k), [=] {
                         return query_case(3000, 700, 3, m, k, 0, 121 + m * 3 + k, CUKD_LIST_HEAP);
                       }});
    cases.push_back({"accumulate method=" + std::to_string(m) + " k=20",
                     [=] { return accumulate_case(2500, 600, 20, m, 131 + m); }});
  }
  // run-time-k lists over Morton-ordered batches of odd sizes
  cases.push_back({"query heap k=20 sorted n=3001", [=] {
                     return query_case(6000, 3001, 3, CUKD_STACK_BASED, 20, CUKD_QUERY_SORT, 151,
                                       CUKD_LIST_HEAP);
                   }});
  cases.push_back({"query fixed k=7 float2 sorted n=2049", [=] {
                     return query_case(5000, 2049, 2, CUKD_STACK_BASED, 7, CUKD_QUERY_SORT, 153);
                   }});
  cases.push_back({"query kNN-8 Morton-sorted", [=] {
                     return query_case(6000, 3000, 3, CUKD_STACK_BASED, 8, CUKD_QUERY_SORT, 47);
                   }});
  // sorted batches of >= 1024 queries run the large-batch kernel in the
  // bigcap build: deep and shallow trees, k = 4 and 16, 2 and 4 dims
  cases.push_back({"query kNN-4 float2 sorted n=4097", [=] {
                     return query_case(4097, 2500, 2, CUKD_STACK_BASED, 4, CUKD_QUERY_SORT, 73);
                   }});
  cases.push_back({"query kNN-16 float4 sorted n=3000", [=] {
                     return query_case(3000, 1500, 4, CUKD_STACK_BASED, 16, CUKD_QUERY_SORT, 79);
                   }});
  cases.push_back({"query FCP sorted n=6000", [=] {
                     return query_case(6000, 3000, 3, CUKD_STACK_BASED, 0, CUKD_QUERY_SORT, 77);
                   }});
  cases.push_back({"query kNN-1 float4 sorted n=5000", [=] {
                     return query_case(5000, 2100, 4, CUKD_STACK_BASED, 1, CUKD_QUERY_SORT, 78);
                   }});
  for (int n : {1, 2, 3, 64})
    cases.push_back({"query kNN-8 sorted tiny n=" + std::to_string(n), [=] {
                       return query_case(n, 1100, 3, CUKD_STACK_BASED, 8, CUKD_QUERY_SORT, 83 + n);
                     }});
  cases.push_back({"query kNN-4 float2 one query per lane", [=] {
                     return query_case(5000, 2000, 2, CUKD_STACK_BASED, 4, CUKD_QUERY_ONE_PER_LANE, 53);
                   }});
  cases.push_back({"query kNN-8 allocation-free sorted", [=] {
                     return query_case(6000, 3000, 3, CUKD_STACK_BASED, 8, CUKD_QUERY_SORT | 64, 59);
                   }});
  cases.push_back({"forest build + kNN-4", [=] {
                     return forest_case({0, 1, 700, 700, 5000, 5003, 12000}, 61);
                   }});
  // finish roots on both sides of level D-1 in one capacity class (the
  // bigcap build puts all three trees at 8192: L0 = 0, 2, 3): a chained
  // and a fixed-order launch over the same workgroups
  cases.push_back({"forest build + kNN-4 mixed finish roots", [=] {
                     return forest_case({0, 3000, 23000, 63000}, 64);
                   }});
  // >= 1024 sorted queries: the large-batch kernel in the bigcap build
  cases.push_back({"forest build + kNN-4 large sorted batch", [=] {
                     return forest_case({0, 1, 700, 700, 5000, 5003, 12000}, 62, 200);
                   }});
  cases.push_back({"spatial float3 build + kNN-8", [=] { return spatial_case(6000, 3, CUKD_SCALAR_F32, 67); }});
  cases.push_back({"spatial int4 build", [=] { return spatial_case(6000, 4, CUKD_SCALAR_I32, 71); }});
  cases.push_back({"spatial float3 kNN-20",
                   [=] { return spatial_case(4000, 3, CUKD_SCALAR_F32, 141, 20); }});
  cases.push_back({"spatial float3 heap kNN-33", [=] {
                     return spatial_case(4000, 3, CUKD_SCALAR_F32, 143, 33, CUKD_LIST_HEAP);
                   }});
  for (auto &c : cases)
    if (only.empty() || c.name.find(only) != std::string::npos) report(c.name, c.run());
  printf(g_fail ? "%d FAILED\n" : "ALL OK\n", g_fail);
  return g_fail ? 1 : 0;
}
