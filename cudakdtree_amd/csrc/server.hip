// server.hip -- a serving loop over a resident forest: requests (a tree, a
// batch of host queries, FCP or kNN-k) arrive at any time from any thread;
// one worker thread batches whatever is queued into one launch per query
// kind (continuous batching), bounded by a device-memory budget, into one
// of SERVER_SLOTS batch slots.  Each slot has a stream of its own and a
// thread that waits for its batch and completes its requests, so batches
// run side by side on the device: a batch held up by a few long traversals
// does not hold up the ones admitted after it.  Inside a batch, light
// requests share one launch per query class and go first; the heavy ones
// (many queries times k) follow, cheapest first, packed into launches of
// bounded work.  Each launch's requests complete as soon as that launch
// has run, so no request waits for much heavier ones it shares a batch
// with, and heavy launches stay large enough to fill the device; a
// request larger than a slot is served in pieces.
// No counterpart in the reference, whose
// queries are per-thread device calls; every result equals the single-tree
// cukd_fcp / cukd_knn result for the same queries.
#include <atomic>
#include <cstdarg>
#include <chrono>
#include <condition_variable>
#include <algorithm>
#include <deque>
#include <mutex>
#include <string.h>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cukd_common.h"

namespace cukd_amd {

namespace {

using Clock = std::chrono::steady_clock;

inline double ms_between(Clock::time_point a, Clock::time_point b)
{
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// what the caller submitted: completed once all of its pieces are
struct Ticket {
  uint64_t id;
  int nq;
  int pieces_left;
  int status;
  bool started;
  Clock::time_point t_submit, t_start;
};

// one piece of a ticket's queries (the whole request unless it is larger
// than a slot)
struct Request {
  Ticket *tk;
  int tree;
  const float *q;
  int nq;
  int k;  // 0 = FCP
  float cutoff;
  int32_t *ids;
  float *dist2;
};

inline size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

// Bytes one (k, cutoff) class of a batch takes in the slot: nq queries in
// nseg segments; the same regions, each 256-byte aligned, that
// launch_batch lays out.
inline size_t class_bytes(size_t nq, size_t nseg, int k, int D)
{
  const size_t kk = k > 0 ? (size_t)k : 1;
  return align_up(nq * D * 4) + align_up((nseg + 1) * 4) + align_up(nseg * 4) +
         2 * align_up(nq * kk * 4);
}

// Work of a request, in candidate-list slots: one with at least HEAVY_COST
// is launched after the light requests of its batch.
inline size_t request_cost(int nq, int k) { return (size_t)nq * (size_t)(k > 0 ? k : 1); }

// HEAVY_GROUP_COST: most work (list slots) packed into one launch of heavy
// requests, unless a single request is larger: about five of the largest
// (100K queries, k = 16).  In a burst over 512 trees, one launch per heavy
// request reached under 2x the per-request throughput and launches of at
// most 2^21 slots 2.3-2.8x; at 2000 requests/s batches hold about one
// request, so the bound matters for bursts only.
enum { SERVER_SLOTS = 8, HEAVY_COST = 1 << 16, HEAVY_GROUP_COST = 1 << 23 };

// one batch: free -> packed by the worker -> launched -> completed by the
// slot's own thread -> free (state guarded by the server's mutex)
enum SlotState { SLOT_FREE = 0, SLOT_PACKING = 1, SLOT_LAUNCHED = 2 };

struct Slot {
  uint8_t *dev = nullptr;   // queries | segment tables | results
  uint8_t *host = nullptr;  // pinned mirror
  size_t bytes = 0;
  hipStream_t stream = nullptr;  // copy-in, then per group: launch, results back
  std::vector<hipEvent_t> cev;   // per group: its results are in 'host'
  struct Group {
    size_t res_off = 0, res_bytes = 0;  // the group's results in the slot
    std::vector<int> members;           // its requests (indices into reqs)
  };
  std::vector<Group> groups;  // in launch order
  std::vector<Request *> reqs;
  std::vector<size_t> ids_off, d2_off;  // per request: where its results are in the slot
  size_t res_begin = 0, used = 0;  // results: [res_begin, used) of the slot
  int state = SLOT_FREE;
  std::thread finisher;
};

// Helper threads for a batch's host-side copies (queries into the pinned
// slot, results out to the callers' buffers): one thread copies at a small
// fraction of the host's memory bandwidth, and these copies bound the
// server's throughput.  run() splits the copies into chunks that the
// helpers and the calling thread take in turn; it returns when every chunk
// is copied and no helper is still inside the generation.  Calls from
// several threads take turns.
class CopyPool {
 public:
  struct Job {
    void *dst;
    const void *src;
    size_t bytes;
  };
  explicit CopyPool(int helpers)
  {
    for (int i = 0; i < helpers; i++) th_.emplace_back([this] { helper(); });
  }
  ~CopyPool()
  {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void run(const std::vector<Job> &jobs)
  {
    enum { CHUNK = 1 << 20, SERIAL = 256 << 10 };
    size_t total = 0;
    for (const Job &j : jobs) total += j.bytes;
    if (th_.empty() || total < SERIAL) {
      for (const Job &j : jobs) memcpy(j.dst, j.src, j.bytes);
      return;
    }
    std::lock_guard<std::mutex> turn(call_mu_);
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [&] { return busy_ == 0; });
    chunks_.clear();
    for (const Job &j : jobs)
      for (size_t o = 0; o < j.bytes; o += CHUNK)
        chunks_.push_back(Job{(uint8_t *)j.dst + o, (const uint8_t *)j.src + o,
                              j.bytes - o < (size_t)CHUNK ? j.bytes - o : (size_t)CHUNK});
    n_ = chunks_.size();
    next_.store(0);
    done_.store(0);
    gen_++;
    lk.unlock();
    cv_.notify_all();
    drain();
    lk.lock();
    cv_done_.wait(lk, [&] { return done_.load() == n_ && busy_ == 0; });
  }

 private:
  void drain()
  {
    for (;;) {
      const size_t i = next_.fetch_add(1);
      if (i >= n_) break;
      memcpy(chunks_[i].dst, chunks_[i].src, chunks_[i].bytes);
      done_.fetch_add(1);
    }
  }
  void helper()
  {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      busy_++;
      lk.unlock();
      drain();
      lk.lock();
      busy_--;
      if (busy_ == 0) cv_done_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex call_mu_;  // one run() at a time
  std::mutex mu_;
  std::condition_variable cv_, cv_done_;
  std::vector<Job> chunks_;
  size_t n_ = 0;
  std::atomic<size_t> next_{0}, done_{0};
  int busy_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

int copy_helpers()
{
  const unsigned hc = std::thread::hardware_concurrency();
  const int h = hc > 2 ? (int)(hc / 2) - 1 : 0;
  return h < 7 ? h : 7;
}

}  // namespace

}  // namespace cukd_amd

using namespace cukd_amd;

struct cukd_server {
  const cukd_forest *forest = nullptr;
  const void *points = nullptr;
  const void *world_bounds = nullptr;
  int traversal = 0;
  int D = 0;
  int max_per_batch = 0;
  int device = 0;
  size_t slot_bytes = 0;
  Slot slot[SERVER_SLOTS];
  std::mutex mu;
  // cv_work: the worker (requests queued and a slot free, or stop);
  // cv_slot: the slots' threads (a batch launched, or finish); cv_done:
  // pollers and the worker's final drain
  std::condition_variable cv_work, cv_slot, cv_done;
  bool finish = false;  // the slots' threads exit once their slot is free
  std::deque<Request *> pending;
  std::deque<cukd_completion> done;
  uint64_t next_ticket = 1;
  bool stop = false;
  int error = 0;
  char error_msg[256] = {0};
  cukd_server_stats stats{};
  CopyPool *copies = nullptr;
  std::thread worker;
};

namespace {

// The launch groups of a batch: the light requests of one (k, cutoff)
// class share a group, in order of first appearance (admission order inside
// it).  The heavy requests of a class, by ascending cost, fill groups of at
// most HEAVY_GROUP_COST.  Light groups go first, then the heavy ones by
// ascending cost.
struct GroupKey {
  int k;
  float cutoff;
  bool heavy;
  size_t cost;  // heavy: the group's cost so far
  bool same(const Request *r, bool h) const
  {
    return h == heavy && k == r->k && memcmp(&cutoff, &r->cutoff, 4) == 0;
  }
};

// Packs the requests of one batch into a slot and launches it: one H2D
// copy, then per group one forest query launch, the copy of its results
// back and an event, so that the slot's thread can complete the group's
// requests as soon as their own launch has run.
int launch_batch(cukd_server *S, Slot &sl)
{
  const int D = S->D;
  std::vector<GroupKey> key;
  std::vector<std::vector<int>> mem;
  std::vector<int> heavy;
  for (int i = 0; i < (int)sl.reqs.size(); i++) {
    const Request *r = sl.reqs[i];
    if (request_cost(r->nq, r->k) >= (size_t)HEAVY_COST) {
      heavy.push_back(i);
      continue;
    }
    size_t g = 0;
    for (; g < key.size(); g++)
      if (key[g].same(r, false)) break;
    if (g == key.size()) {
      key.push_back(GroupKey{r->k, r->cutoff, false, 0});
      mem.emplace_back();
    }
    mem[g].push_back(i);
  }
  std::stable_sort(heavy.begin(), heavy.end(), [&](int x, int y) {
    return request_cost(sl.reqs[x]->nq, sl.reqs[x]->k) < request_cost(sl.reqs[y]->nq, sl.reqs[y]->k);
  });
  const size_t light_groups = key.size();
  for (int i : heavy) {
    const Request *r = sl.reqs[i];
    const size_t c = request_cost(r->nq, r->k);
    // the class's open group (the last one made), if the request still fits
    size_t g = key.size();
    for (size_t j = key.size(); j > light_groups; j--)
      if (key[j - 1].same(r, true)) {
        if (key[j - 1].cost + c <= (size_t)HEAVY_GROUP_COST) g = j - 1;
        break;
      }
    if (g == key.size()) {
      key.push_back(GroupKey{r->k, r->cutoff, true, 0});
      mem.emplace_back();
    }
    key[g].cost += c;
    mem[g].push_back(i);
  }
  std::vector<int> order(key.size());
  for (size_t g = 0; g < key.size(); g++) order[g] = (int)g;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (key[x].heavy != key[y].heavy) return !key[x].heavy;
    return key[x].heavy && key[x].cost < key[y].cost;
  });
  // layout: per group [queries][seg_off][seg_tree]; then per group results
  struct Lay { size_t q, so, st, ids, d2; int nq, nseg; };
  const size_t G = key.size();
  std::vector<Lay> lay(G);
  size_t o = 0;
  for (size_t j = 0; j < G; j++) {
    const int g = order[j];
    int nq = 0;
    for (int i : mem[g]) nq += sl.reqs[i]->nq;
    lay[j].nq = nq;
    lay[j].nseg = (int)mem[g].size();
    lay[j].q = o; o = align_up(o + (size_t)nq * D * 4);
    lay[j].so = o; o = align_up(o + (size_t)(lay[j].nseg + 1) * 4);
    lay[j].st = o; o = align_up(o + (size_t)lay[j].nseg * 4);
  }
  const size_t res_begin = o;
  sl.ids_off.assign(sl.reqs.size(), 0);
  sl.d2_off.assign(sl.reqs.size(), 0);
  sl.groups.assign(G, Slot::Group());
  for (size_t j = 0; j < G; j++) {
    const size_t kk = key[order[j]].k > 0 ? (size_t)key[order[j]].k : 1;
    lay[j].ids = o; o = align_up(o + (size_t)lay[j].nq * kk * 4);
    lay[j].d2 = o; o = align_up(o + (size_t)lay[j].nq * kk * 4);
    sl.groups[j].res_off = lay[j].ids;
    sl.groups[j].res_bytes = o - lay[j].ids;
    sl.groups[j].members = mem[order[j]];
  }
  if (o > sl.bytes) {
    // admission charges at least this layout (every heavy request as a
    // group of its own), so this is not reached
    set_last_error("batch of %zu bytes exceeds the slot of %zu", o, sl.bytes);
    return CUKD_ERR_INVALID;
  }
  hipError_t e = hipSuccess;
  while (sl.cev.size() < G && e == hipSuccess) {
    hipEvent_t c = nullptr;
    e = hipEventCreateWithFlags(&c, hipEventDisableTiming);
    if (c) sl.cev.push_back(c);
  }
  if (e != hipSuccess || sl.cev.size() < G) {
    set_last_error("batch events: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  sl.res_begin = res_begin;
  sl.used = o;
  std::vector<CopyPool::Job> jobs;
  for (size_t j = 0; j < G; j++) {
    float *hq = (float *)(sl.host + lay[j].q);
    int32_t *so = (int32_t *)(sl.host + lay[j].so);
    int32_t *stt = (int32_t *)(sl.host + lay[j].st);
    int acc = 0, sg = 0;
    const size_t kk = key[order[j]].k > 0 ? (size_t)key[order[j]].k : 1;
    for (int i : sl.groups[j].members) {
      const Request *r = sl.reqs[i];
      jobs.push_back(CopyPool::Job{hq + (size_t)acc * D, r->q, (size_t)r->nq * D * 4});
      so[sg] = acc;
      stt[sg] = r->tree;
      sl.ids_off[i] = lay[j].ids + (size_t)acc * kk * 4;
      sl.d2_off[i] = lay[j].d2 + (size_t)acc * kk * 4;
      acc += r->nq;
      sg++;
    }
    so[sg] = acc;
  }
  S->copies->run(jobs);
  e = hipMemcpyAsync(sl.dev, sl.host, res_begin, hipMemcpyHostToDevice, sl.stream);
  for (size_t j = 0; j < G && e == hipSuccess; j++) {
    const GroupKey &gk = key[order[j]];
    const float *dq = (const float *)(sl.dev + lay[j].q);
    const int32_t *dso = (const int32_t *)(sl.dev + lay[j].so);
    const int32_t *dst = (const int32_t *)(sl.dev + lay[j].st);
    int32_t *ids = (int32_t *)(sl.dev + lay[j].ids);
    float *d2 = (float *)(sl.dev + lay[j].d2);
    int rc;
    if (gk.k == 0)
      rc = cukd_forest_fcp(S->traversal, S->forest, S->points, S->world_bounds, dq, lay[j].nq,
                           lay[j].nseg, dso, dst, gk.cutoff, 0.f, ids, d2, 0, sl.stream);
    else
      rc = cukd_forest_knn(S->traversal, CUKD_LIST_FIXED, gk.k, S->forest, S->points,
                           S->world_bounds, dq, lay[j].nq, lay[j].nseg, dso, dst, gk.cutoff,
                           ids, d2, nullptr, 0, sl.stream);
    if (rc != CUKD_OK) return rc;
    // the group's results come back right after its own launch, before the
    // next group's; the slot's thread completes the group's requests once
    // this copy is done
    e = hipMemcpyAsync(sl.host + sl.groups[j].res_off, sl.dev + sl.groups[j].res_off,
                       sl.groups[j].res_bytes, hipMemcpyDeviceToHost, sl.stream);
    if (e == hipSuccess) e = hipEventRecord(sl.cev[j], sl.stream);
  }
  if (e != hipSuccess) {
    set_last_error("batch copy: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

// Completes one piece (the server's mutex held): its ticket completes with
// its last piece.
void complete_piece(cukd_server *S, Request *r, int status, Clock::time_point now);

// Waits for each group of a slot's batch in launch order, copies its
// results on to its requests' callers and completes them; the slot is
// free once every group is done.
int finish_batch(cukd_server *S, Slot &sl)
{
  hipError_t e = hipSuccess;
  for (size_t j = 0; j < sl.groups.size(); j++) {
    const Slot::Group &gr = sl.groups[j];
    if (e == hipSuccess) e = hipEventSynchronize(sl.cev[j]);
    if (e == hipSuccess) {
      std::vector<CopyPool::Job> jobs;
      for (int i : gr.members) {
        const Request *r = sl.reqs[i];
        const size_t kk = r->k > 0 ? (size_t)r->k : 1;
        jobs.push_back(CopyPool::Job{r->ids, sl.host + sl.ids_off[i], (size_t)r->nq * kk * 4});
        if (r->dist2)
          jobs.push_back(CopyPool::Job{r->dist2, sl.host + sl.d2_off[i], (size_t)r->nq * kk * 4});
      }
      S->copies->run(jobs);
    }
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    for (int i : gr.members) {
      complete_piece(S, sl.reqs[i], e == hipSuccess ? CUKD_OK : CUKD_ERR_HIP, now);
      sl.reqs[i] = nullptr;
    }
    S->cv_done.notify_all();
  }
  std::lock_guard<std::mutex> lk(S->mu);
  sl.reqs.clear();
  sl.state = SLOT_FREE;
  S->cv_work.notify_one();
  if (e != hipSuccess && !S->error) {
    S->error = CUKD_ERR_HIP;
    snprintf(S->error_msg, sizeof(S->error_msg), "batch failed: %s", hipGetErrorString(e));
  }
  S->cv_done.notify_all();
  return e == hipSuccess ? CUKD_OK : CUKD_ERR_HIP;
}

void complete_piece(cukd_server *S, Request *r, int status, Clock::time_point now)
{
  Ticket *tk = r->tk;
  if (status != CUKD_OK && tk->status == CUKD_OK) tk->status = status;
  S->stats.queries_done += (uint64_t)r->nq;
  delete r;
  if (--tk->pieces_left > 0) return;
  cukd_completion c;
  c.ticket = tk->id;
  c.status = tk->status;
  c.num_queries = tk->nq;
  c.queue_ms = (float)ms_between(tk->t_submit, tk->t_start);
  c.total_ms = (float)ms_between(tk->t_submit, now);
  S->done.push_back(c);
  S->stats.requests_done++;
  delete tk;
}

// The worker: admits queued requests into a free slot and launches the
// batch on the slot's stream; the slot's own thread completes it.
void worker_main(cukd_server *S)
{
  (void)hipSetDevice(S->device);
  for (;;) {
    int si = -1;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      auto free_slot = [&] {
        for (int i = 0; i < SERVER_SLOTS; i++)
          if (S->slot[i].state == SLOT_FREE) return i;
        return -1;
      };
      S->cv_work.wait(lk, [&] {
        return (S->stop && S->pending.empty()) || (!S->pending.empty() && free_slot() >= 0);
      });
      if (S->pending.empty()) break;  // stopping, and every request admitted
      si = free_slot();
      Slot &sl = S->slot[si];
      // admit requests in arrival order while the batch fits the slot, laid
      // out as launch_batch lays it out (one region set per launch group),
      // with every heavy request charged as a group of its own: packing
      // heavy requests together only removes padding
      struct Cls { GroupKey key; size_t nq, nseg, bytes; };
      std::vector<Cls> cls;
      size_t bytes = 0;
      const Clock::time_point now = Clock::now();
      while (!S->pending.empty() && (S->max_per_batch <= 0 || (int)sl.reqs.size() < S->max_per_batch)) {
        Request *r = S->pending.front();
        const bool heavy = request_cost(r->nq, r->k) >= (size_t)HEAVY_COST;
        size_t c = 0;
        for (; c < cls.size() && !heavy; c++)
          if (cls[c].key.same(r, false)) break;
        if (heavy) c = cls.size();
        const size_t old_b = c < cls.size() ? cls[c].bytes : 0;
        const size_t nq = (c < cls.size() ? cls[c].nq : 0) + (size_t)r->nq;
        const size_t nseg = (c < cls.size() ? cls[c].nseg : 0) + 1;
        const size_t new_b = class_bytes(nq, nseg, r->k, S->D);
        if (!sl.reqs.empty() && bytes - old_b + new_b > S->slot_bytes) break;
        if (c == cls.size()) cls.push_back(Cls{GroupKey{r->k, r->cutoff, heavy, 0}, 0, 0, 0});
        cls[c].nq = nq;
        cls[c].nseg = nseg;
        cls[c].bytes = new_b;
        bytes = bytes - old_b + new_b;
        if (!r->tk->started) {
          r->tk->started = true;
          r->tk->t_start = now;
        }
        sl.reqs.push_back(r);
        S->pending.pop_front();
      }
      sl.state = SLOT_PACKING;
    }
    Slot &sl = S->slot[si];
    const int rc = launch_batch(S, sl);
    std::lock_guard<std::mutex> lk(S->mu);
    if (rc != CUKD_OK) {
      // fail the batch's requests; only a failed HIP call marks the server
      // itself as failed
      const Clock::time_point now = Clock::now();
      for (Request *r : sl.reqs) complete_piece(S, r, rc, now);
      sl.reqs.clear();
      sl.state = SLOT_FREE;
      if (rc == CUKD_ERR_HIP && !S->error) {
        S->error = CUKD_ERR_HIP;
        snprintf(S->error_msg, sizeof(S->error_msg), "batch launch: %s", cukd_last_error());
      }
      S->cv_done.notify_all();
      continue;
    }
    sl.state = SLOT_LAUNCHED;
    S->stats.batches++;
    S->cv_slot.notify_all();
  }
  // every admitted batch completes before the slots' threads are released
  std::unique_lock<std::mutex> lk(S->mu);
  S->cv_work.wait(lk, [&] {
    for (int i = 0; i < SERVER_SLOTS; i++)
      if (S->slot[i].state != SLOT_FREE) return false;
    return true;
  });
  S->finish = true;
  S->cv_slot.notify_all();
}

// A slot's thread: waits for each batch launched into the slot, copies its
// results to the callers and records the completions.
void slot_main(cukd_server *S, int si)
{
  (void)hipSetDevice(S->device);
  Slot &sl = S->slot[si];
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_slot.wait(lk, [&] { return sl.state == SLOT_LAUNCHED || S->finish; });
      if (sl.state != SLOT_LAUNCHED) return;
    }
    finish_batch(S, sl);
  }
}

}  // namespace

extern "C" {

int cukd_server_create(cukd_server **server, const cukd_forest *forest, const void *d_points,
                       const void *d_world_bounds, int traversal, size_t budget_bytes,
                       int32_t max_requests_per_batch, hipStream_t stream)
{
  clear_last_error();
  if (!server || !forest) { set_last_error("null argument"); return CUKD_ERR_INVALID; }
  *server = nullptr;
  if (traversal < CUKD_STACK_BASED || traversal > CUKD_CCT) {
    set_last_error("unknown traversal %d", traversal);
    return CUKD_ERR_INVALID;
  }
  if (traversal == CUKD_CCT && !d_world_bounds) {
    set_last_error("cct traversal needs world bounds");
    return CUKD_ERR_INVALID;
  }
  if (budget_bytes < (1u << 20)) { set_last_error("budget below 1 MiB"); return CUKD_ERR_INVALID; }
  cukd_server *S = new cukd_server();
  S->forest = forest;
  S->points = d_points;
  S->world_bounds = d_world_bounds;
  S->traversal = traversal;
  S->max_per_batch = max_requests_per_batch;
  S->D = forest_num_dims(forest);
  if (hipGetDevice(&S->device) != hipSuccess) S->device = 0;
  S->slot_bytes = budget_bytes / SERVER_SLOTS;
  // the slots' streams start after the work queued on 'stream' so far
  // (the forest's build, typically)
  hipEvent_t ready = nullptr;
  hipError_t e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ready, stream);
  for (int i = 0; i < SERVER_SLOTS && e == hipSuccess; i++) {
    Slot &sl = S->slot[i];
    sl.bytes = S->slot_bytes;
    e = hipMalloc((void **)&sl.dev, sl.bytes);
    if (e == hipSuccess) e = hipHostMalloc((void **)&sl.host, sl.bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamWaitEvent(sl.stream, ready, 0);
  }
  if (ready) (void)hipEventDestroy(ready);
  if (e != hipSuccess) {
    set_last_error("server set-up failed: %s", hipGetErrorString(e));
    for (int j = 0; j < SERVER_SLOTS; j++) {
      Slot &sl = S->slot[j];
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
      if (sl.dev) (void)hipFree(sl.dev);
      if (sl.host) (void)hipHostFree(sl.host);
    }
    delete S;
    return CUKD_ERR_ALLOC;
  }
  S->copies = new CopyPool(copy_helpers());
  for (int i = 0; i < SERVER_SLOTS; i++) S->slot[i].finisher = std::thread(slot_main, S, i);
  S->worker = std::thread(worker_main, S);
  *server = S;
  return CUKD_OK;
}

int cukd_server_submit(cukd_server *S, int32_t tree, const float *h_queries, int32_t nq, int k,
                       float cutoff, int32_t *h_ids, float *h_dist2, uint64_t *ticket)
{
  clear_last_error();
  if (!S || !ticket) { set_last_error("null argument"); return CUKD_ERR_INVALID; }
  if (tree < 0 || tree >= forest_num_trees(S->forest)) {
    set_last_error("tree %d out of range", tree);
    return CUKD_ERR_INVALID;
  }
  if (nq < 1 || !h_queries || !h_ids) { set_last_error("empty request"); return CUKD_ERR_INVALID; }
  if (k < 0 || k > 64) { set_last_error("k = %d not in 0..64", k); return CUKD_ERR_INVALID; }
  if (k > 0 && !h_dist2) { set_last_error("kNN needs h_dist2"); return CUKD_ERR_INVALID; }
  // pieces of at most 'piece' queries, each of which fits a slot alone
  int piece = nq;
  while (piece > 1 && class_bytes((size_t)piece, 1, k, S->D) > S->slot_bytes) piece = (piece + 1) / 2;
  if (class_bytes((size_t)piece, 1, k, S->D) > S->slot_bytes) {
    set_last_error("the server's memory budget cannot hold even one query");
    return CUKD_ERR_INVALID;
  }
  const int pieces = (nq + piece - 1) / piece;
  Ticket *tk = new Ticket();
  tk->nq = nq;
  tk->pieces_left = pieces;
  tk->status = CUKD_OK;
  tk->started = false;
  tk->t_submit = Clock::now();
  const size_t kk = k > 0 ? (size_t)k : 1;
  std::vector<Request *> rs((size_t)pieces);
  for (int p = 0; p < pieces; p++) {
    const int a = p * piece, n = nq - a < piece ? nq - a : piece;
    Request *r = new Request();
    r->tk = tk;
    r->tree = tree;
    r->q = h_queries + (size_t)a * S->D;
    r->nq = n;
    r->k = k;
    r->cutoff = cutoff;
    r->ids = h_ids + (size_t)a * kk;
    r->dist2 = h_dist2 ? h_dist2 + (size_t)a * kk : nullptr;
    rs[p] = r;
  }
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) {
      for (Request *r : rs) delete r;
      delete tk;
      set_last_error("server is shutting down");
      return CUKD_ERR_INVALID;
    }
    tk->id = S->next_ticket++;
    *ticket = tk->id;
    for (Request *r : rs) S->pending.push_back(r);
    S->stats.requests_submitted++;
  }
  S->cv_work.notify_one();
  return CUKD_OK;
}

int cukd_server_poll(cukd_server *S, cukd_completion *out, int32_t max, int32_t *count,
                     int32_t wait_for)
{
  clear_last_error();
  if (!S || !count || (max > 0 && !out)) { set_last_error("null argument"); return CUKD_ERR_INVALID; }
  std::unique_lock<std::mutex> lk(S->mu);
  if (wait_for > 0)
    S->cv_done.wait(lk, [&] { return (int)S->done.size() >= wait_for || S->error; });
  int n = 0;
  while (n < max && !S->done.empty()) {
    out[n++] = S->done.front();
    S->done.pop_front();
  }
  *count = n;
  if (S->error) {
    set_last_error("%s", S->error_msg);
    return S->error;
  }
  return CUKD_OK;
}

int cukd_server_get_stats(cukd_server *S, cukd_server_stats *stats)
{
  if (!S || !stats) return CUKD_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  *stats = S->stats;
  return CUKD_OK;
}

int cukd_server_destroy(cukd_server *S)
{
  clear_last_error();
  if (!S) return CUKD_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  S->worker.join();  // serves every pending request first
  for (int i = 0; i < SERVER_SLOTS; i++) S->slot[i].finisher.join();
  delete S->copies;
  for (int i = 0; i < SERVER_SLOTS; i++) {
    Slot &sl = S->slot[i];
    (void)hipStreamDestroy(sl.stream);
    (void)hipFree(sl.dev);
    (void)hipHostFree(sl.host);
    for (hipEvent_t ev : sl.cev) (void)hipEventDestroy(ev);
  }
  const int err = S->error;
  delete S;
  return err;
}

}  // extern "C"
