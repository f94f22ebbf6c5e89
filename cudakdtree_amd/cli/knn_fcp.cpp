// knn_fcp.cpp -- command-line driver with the reference test driver's
// interface (testing/floatN-knn-and-fcp.cu:514-558, measure.sh:13-52):
//
//   cukd_knn_fcp [N] [-v] [-nq numQueries] [-nr repeats] [--clustered]
//                [-r cutOffRadius] [-k k] [-lt leafThreshold]
//                [--load-dumped-files]
//                [--dims 2|3|4|8] [--knn|--fcp] [--xd] [--stats] [--query-sort]
//                [--method stackBased|stackFree|cct|spatial-stackBased|spatial-cct]
//                [--dump-files]
//
// The reference compiles one binary per (dims, fcp/knn, method, -xd); here
// those are run-time flags.  Points and queries come from the same
// generators as the reference driver (uniform integers in [0, 1e6], or the
// clustered variant), or with --load-dumped-files from the files
// "data_points" / "query_points" in the point-file format (size_t count,
// then raw floatN rows); --dump-files writes the generated sets there.
// Output lines match the reference driver's ("done building tree, took",
// "that is ... queries/s", "CHECKSUM", "NICE_STATS"), so measure.sh-style
// sweeps work unchanged.  -v checks the tree and every query result
// against a host brute force with the reference's tolerance rule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <queue>
#include <random>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "cukd_amd.h"

namespace {

void check(int rc, const char *what)
{
  if (rc != CUKD_OK) throw std::runtime_error(std::string(what) + ": " + cukd_last_error());
}

void hip_check(hipError_t e, const char *what)
{
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

double now_s()
{
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// common.h prettyDouble: one decimal and an SI suffix
std::string pretty(double v)
{
  const double a = std::fabs(v);
  char b[64];
  struct { double s; char c; } up[] = {{1e18, 'E'}, {1e15, 'P'}, {1e12, 'T'}, {1e9, 'G'}, {1e6, 'M'}, {1e3, 'k'}};
  for (auto &u : up)
    if (a >= u.s) { snprintf(b, sizeof(b), "%.1f%c", (float)(v / u.s), u.c); return b; }
  struct { double s; char c; } dn[] = {{1e-12, 'f'}, {1e-9, 'p'}, {1e-6, 'n'}, {1e-3, 'u'}, {1.0, 'm'}};
  const double mul[] = {1e15, 1e12, 1e9, 1e6, 1e3};
  for (int i = 0; i < 5; i++)
    if (a <= dn[i].s) { snprintf(b, sizeof(b), "%.1f%c", (float)(v * mul[i]), dn[i].c); return b; }
  snprintf(b, sizeof(b), "%f", (float)v);
  return b;
}

// the reference driver's generators: mt19937 seeded through a seed_seq and
// default_random_engine, integer coordinates in [0, 1e6]; one seed per
// call, counting up from 100000 (uniform) or 3214567 (clustered)
std::vector<float> generate_uniform(int N, int D)
{
  static int seed = 100000;
  std::seed_seq seq{seed++};
  std::default_random_engine rd(seq);
  std::mt19937 gen(rd());
  std::uniform_int_distribution<> dist(0, 1000000);
  std::vector<float> p((size_t)N * D);
  for (size_t i = 0; i < p.size(); i++) p[i] = (float)dist(gen);
  return p;
}

std::vector<float> generate_clustered(int N, int D)
{
  static int seed = 3214567;
  std::seed_seq seq{seed++};
  const int nc = std::max(1, (int)sqrtf((float)N));
  std::default_random_engine rd(seq);
  std::mt19937 gen(rd());
  std::uniform_int_distribution<> dist(0, 1000000);
  // cluster c sits in cell (c, perm_1(c), perm_2(c), ...) of an nc^D grid
  std::vector<int> cell((size_t)nc * D);
  for (int c = 0; c < nc; c++)
    for (int d = 0; d < D; d++) cell[(size_t)c * D + d] = c;
  for (int d = 1; d < D; d++)
    for (int i = nc - 1; i > 0; --i) {
      const int j = dist(gen) % i;
      if (i != j) std::swap(cell[(size_t)i * D + d], cell[(size_t)j * D + d]);
    }
  std::vector<float> p((size_t)N * D);
  for (int i = 0; i < N; i++) {
    const int c = i % nc;
    for (int d = 0; d < D; d++) {
      const long long cd = cell[(size_t)c * D + d];
      const int b = (int)(cd * N / nc), e = (int)((cd + 1) * N / nc);
      const int r = dist(gen);
      p[(size_t)i * D + d] = (float)(b + r % std::max(1, e - b));
    }
  }
  return p;
}

std::vector<float> load_file(const char *name, int D, size_t &count)
{
  std::ifstream in(name, std::ios::binary);
  if (!in) throw std::runtime_error(std::string("cannot open ") + name);
  in.read((char *)&count, sizeof(count));
  std::vector<float> p(count * D);
  in.read((char *)p.data(), (std::streamsize)(p.size() * 4));
  if (!in) throw std::runtime_error(std::string(name) + " is shorter than its count");
  std::cout << "loaded " << count << " points from " << name << std::endl;
  return p;
}

void save_file(const char *name, const std::vector<float> &p, int D)
{
  std::ofstream out(name, std::ios::binary);
  const size_t count = p.size() / D;
  out.write((const char *)&count, sizeof(count));
  out.write((const char *)p.data(), (std::streamsize)(p.size() * 4));
}

// host squared distance, evaluated left to right without contraction like
// the library's kernels (cukd-math.h:210-213)
#pragma STDC FP_CONTRACT OFF
inline float sqr_dist(const float *a, const float *b, int D)
{
  float s = 0.f;
  for (int d = 0; d < D; d++) {
    const float t = a[d] - b[d];
    const float t2 = t * t;
    s = d == 0 ? t2 : s + t2;
  }
  return s;
}

bool close_enough(float actual, float reported)
{
  return actual == reported ||
         std::fabs(actual - reported) / std::max(actual, reported) <= 1e-6f;
}

template <class F>
void parallel_for(size_t n, F f)
{
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      for (size_t i = t; i < n; i += nt) f(i);
    });
  for (auto &x : th) x.join();
}

}  // namespace

int main(int ac, const char **av)
{
  try {
    int N = 1000000, nq = 1000000, nr = 1, k = 50, D = 3, lt = 0;
    bool verify = false, clustered = false, knn = false, xd = false, stats = false, sorted = false;
    bool load = false, dump = false;
    float cutoff = std::numeric_limits<float>::infinity();
    std::string method = "stackBased";
    for (int i = 1; i < ac; i++) {
      const std::string a = av[i];
      auto next = [&]() -> const char * {
        if (i + 1 >= ac) throw std::runtime_error("missing value after " + a);
        return av[++i];
      };
      if (a[0] != '-') N = std::stoi(a);
      else if (a == "-v") verify = true;
      else if (a == "-nq") nq = std::stoi(next());
      else if (a == "-nr") nr = std::stoi(next());
      else if (a == "--clustered") clustered = true;
      else if (a == "-r") cutoff = std::stof(next());
      else if (a == "-k") { k = std::stoi(next()); knn = true; }
      else if (a == "-lt") lt = std::stoi(next());
      else if (a == "--load-dumped-files") load = true;
      else if (a == "--dump-files") dump = true;
      else if (a == "--dims") D = std::stoi(next());
      else if (a == "--knn") knn = true;
      else if (a == "--fcp") knn = false;
      else if (a == "--xd") xd = true;
      else if (a == "--stats") stats = true;
      else if (a == "--query-sort") sorted = true;
      else if (a == "--method") method = next();
      else throw std::runtime_error("unknown cmdline arg " + a);
    }
    const bool spatial = method.rfind("spatial-", 0) == 0;
    const std::string trav_name = spatial ? method.substr(8) : method;
    int trav;
    if (trav_name == "stackBased") trav = CUKD_STACK_BASED;
    else if (trav_name == "stackFree" && !spatial) trav = CUKD_STACK_FREE;
    else if (trav_name == "cct") trav = CUKD_CCT;
    else throw std::runtime_error("unknown method " + method);
    if (D != 2 && D != 3 && D != 4 && D != 8) throw std::runtime_error("--dims must be 2, 3, 4 or 8");
    if (spatial && xd) throw std::runtime_error("--xd applies to the left-balanced tree only");

    std::vector<float> pts;
    size_t npts = (size_t)N;
    if (load) pts = load_file("data_points", D, npts);
    else {
      std::cout << "generating " << N << (clustered ? " clustered" : " uniform random") << " points" << std::endl;
      pts = clustered ? generate_clustered(N, D) : generate_uniform(N, D);
    }
    N = (int)npts;
    std::vector<float> qs;
    size_t nqs = (size_t)nq;
    if (load) qs = load_file("query_points", D, nqs);
    else qs = generate_uniform(nq, D);
    nq = (int)nqs;
    if (dump) {
      save_file("data_points", pts, D);
      save_file("query_points", qs, D);
    }

    // element layout: floatN, or PointAndDim {floatN point; int dim} (-xd)
    cukd_layout L{D, CUKD_SCALAR_F32, 4 * D, 0, -1, 4};
    if (xd) { L.stride = 4 * D + 4; L.dim_offset = 4 * D; }
    const int W = L.stride / 4;
    std::vector<float> elems((size_t)N * W, 0.f);
    for (int i = 0; i < N; i++)
      for (int d = 0; d < D; d++) elems[(size_t)i * W + d] = pts[(size_t)i * D + d];

    void *d_pts = nullptr, *d_q = nullptr, *d_bounds = nullptr;
    int32_t *d_ids = nullptr;
    float *d_d2 = nullptr, *d_r = nullptr;
    unsigned long long *d_stats = nullptr;
    const int kk = knn ? k : 1;
    hip_check(hipMalloc(&d_pts, std::max<size_t>(1, elems.size() * 4)), "hipMalloc");
    hip_check(hipMalloc(&d_q, std::max<size_t>(1, qs.size() * 4)), "hipMalloc");
    hip_check(hipMalloc(&d_bounds, 2 * D * 4), "hipMalloc");
    hip_check(hipMalloc((void **)&d_ids, std::max<size_t>(1, (size_t)nq * kk * 4)), "hipMalloc");
    hip_check(hipMalloc((void **)&d_d2, std::max<size_t>(1, (size_t)nq * kk * 4)), "hipMalloc");
    hip_check(hipMalloc((void **)&d_r, std::max<size_t>(1, (size_t)nq * 4)), "hipMalloc");
    hip_check(hipMalloc((void **)&d_stats, 8), "hipMalloc");
    hip_check(hipMemcpy(d_pts, elems.data(), elems.size() * 4, hipMemcpyHostToDevice), "copy");
    hip_check(hipMemcpy(d_q, qs.data(), qs.size() * 4, hipMemcpyHostToDevice), "copy");

    cukd_spatial_tree st{};
    std::cout << "calling builder..." << std::endl;
    double t0 = now_s();
    if (spatial) check(cukd_spatial_build(&st, d_pts, N, &L, lt, 0, nullptr), "spatial build");
    else check(cukd_build_tree(d_pts, N, &L, CUKD_BUILDER_ID_DEFAULT, d_bounds, 0, nullptr), "build");
    hip_check(hipDeviceSynchronize(), "build");
    double t1 = now_s();
    std::cout << "done building tree, took " << pretty(t1 - t0) << "s" << std::endl;

    std::vector<float> tree((size_t)N * W);
    hip_check(hipMemcpy(tree.data(), d_pts, tree.size() * 4, hipMemcpyDeviceToHost), "copy");
    if (verify && !spatial) {
      // every node on the correct side of each ancestor's split plane
      // (floatN-knn-and-fcp.cu:387-427)
      for (int i = 1; i < N; i++)
        for (int c = i; c > 0; c = (c - 1) / 2) {
          const int a = (c - 1) / 2;
          int dim;
          if (xd) memcpy(&dim, &tree[(size_t)a * W + D], 4);
          else dim = (31 - __builtin_clz((unsigned)(a + 1))) % D;
          const float v = tree[(size_t)i * W + dim], s = tree[(size_t)a * W + dim];
          if ((c & 1) ? v > s : v < s)
            throw std::runtime_error("invalid k-d tree at node " + std::to_string(i));
        }
      std::vector<float> a = pts, b((size_t)N * D);
      for (int i = 0; i < N; i++)
        for (int d = 0; d < D; d++) b[(size_t)i * D + d] = tree[(size_t)i * W + d];
      auto rows = [&](std::vector<float> &v) {
        std::vector<std::vector<float>> r(N);
        for (int i = 0; i < N; i++) r[i].assign(v.begin() + (size_t)i * D, v.begin() + (size_t)(i + 1) * D);
        std::sort(r.begin(), r.end());
        return r;
      };
      if (rows(a) != rows(b)) throw std::runtime_error("tree is not a permutation of the input");
      std::cout << "** verify: tree checked, and valid k-d tree" << std::endl;
    }

    auto run = [&]() {
      if (spatial) {
        if (knn)
          check(cukd_spatial_knn(trav, CUKD_LIST_FIXED, k, &st, &L, (const float *)d_q, nq, cutoff,
                                 d_ids, d_d2, d_r, 0),
                "knn");
        else
          check(cukd_spatial_fcp(trav, &st, &L, (const float *)d_q, nq, cutoff, d_ids, d_d2, 0), "fcp");
      } else if (knn) {
        check(cukd_knn(trav, CUKD_LIST_FIXED, k, d_pts, N, &L, d_bounds, (const float *)d_q, nq,
                       cutoff, d_ids, d_d2, d_r, sorted ? CUKD_QUERY_SORT : 0, 0, nullptr),
              "knn");
      } else {
        check(cukd_fcp(trav, d_pts, N, &L, d_bounds, (const float *)d_q, nq, cutoff, 0.f, d_ids,
                       d_d2, sorted ? CUKD_QUERY_SORT : 0, 0, nullptr),
              "fcp");
      }
    };
    // first call: node-visit statistics (CUKD_STATS semantics)
    hip_check(hipMemset(d_stats, 0, 8), "memset");
    if (stats) check(cukd_set_traversal_stats(d_stats), "stats");
    run();
    hip_check(hipDeviceSynchronize(), "query");
    check(cukd_set_traversal_stats(nullptr), "stats");
    unsigned long long visits = 0;
    hip_check(hipMemcpy(&visits, d_stats, 8, hipMemcpyDeviceToHost), "copy");
    std::cout << "NICE_STATS " << (stats ? pretty((double)visits) : std::string("0")) << std::endl;
    if (stats) std::cout << "KDTREE_STATS " << visits << std::endl;
    t0 = now_s();
    for (int r = 0; r < nr; r++) run();
    hip_check(hipDeviceSynchronize(), "query");
    t1 = now_s();
    std::cout << "done " << nr << " iterations of " << nq << (knn ? " knn " : " fcp ")
              << " queries, took " << pretty(t1 - t0) << "s" << std::endl;
    std::cout << "that is " << pretty((double)nq * nr / (t1 - t0)) << " queries/s" << std::endl;

    // per-query result as the reference driver reports it: distance to the
    // closest point (inf if none), or sqrt of the kNN max radius
    std::vector<float> res(nq);
    {
      std::vector<int32_t> ids((size_t)nq * kk);
      std::vector<float> d2((size_t)nq * kk), r(nq);
      hip_check(hipMemcpy(ids.data(), d_ids, ids.size() * 4, hipMemcpyDeviceToHost), "copy");
      hip_check(hipMemcpy(d2.data(), d_d2, d2.size() * 4, hipMemcpyDeviceToHost), "copy");
      if (knn) hip_check(hipMemcpy(r.data(), d_r, r.size() * 4, hipMemcpyDeviceToHost), "copy");
      for (int i = 0; i < nq; i++)
        res[i] = knn ? sqrtf(r[i]) : (ids[i] < 0 ? INFINITY : sqrtf(d2[i]));
    }
    double checksum = 0.;
    for (int i = 0; i < nq; i++) checksum += res[i];
    std::cout << "CHECKSUM " << checksum << std::endl;

    if (verify) {
      std::cout << "verifying ..." << std::endl;
      std::vector<int> bad(nq, 0);
      parallel_for((size_t)nq, [&](size_t i) {
        const float *q = &qs[i * D];
        const float r2 = cutoff * cutoff;
        float actual;
        if (!knn) {
          actual = INFINITY;
          for (int j = 0; j < N; j++) {
            const float d = sqr_dist(q, &pts[(size_t)j * D], D);
            if (d >= r2) continue;
            actual = std::min(actual, sqrtf(d));
          }
        } else {
          std::priority_queue<float> best;
          for (int j = 0; j < N; j++) {
            const float d = sqr_dist(q, &pts[(size_t)j * D], D);
            if (d >= r2) continue;
            best.push(d);
            if ((int)best.size() > k) best.pop();
          }
          actual = (int)best.size() == k ? sqrtf(best.top()) : cutoff;
        }
        bad[i] = !close_enough(actual, res[i]);
      });
      for (int i = 0; i < nq; i++)
        if (bad[i]) throw std::runtime_error("verification failed for query " + std::to_string(i));
      std::cout << "verification succeeded... done." << std::endl;
    }
    if (spatial) check(cukd_spatial_free(&st, 0, nullptr), "free");
    (void)hipFree(d_pts); (void)hipFree(d_q); (void)hipFree(d_bounds); (void)hipFree(d_ids);
    (void)hipFree(d_d2); (void)hipFree(d_r); (void)hipFree(d_stats);
    return 0;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
