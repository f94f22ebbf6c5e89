// builder.hip -- left-balanced k-d tree builder for gfx950.
//
// Produces exactly the array buildTree_host produces (builder_thrust.h:
// 514-601).  The host builder stable-sorts (tag, coord[dim]) once per
// level, so within a level-L subtree the elements end up ordered by the
// composite key
//     (c[d_L], c[d_{L-1}], ..., c[d_0], input index)
// where d_j is the split dim of the ancestor on level j (repeated dims
// dropped: they never decide a comparison).  The pivot of a subtree is the
// element whose rank under that key equals the left-subtree size
// (ArrayLayoutInStep::pivotPosOf, helpers.h:235-241).  The tree depends
// only on those selections, so the GPU builder never materialises the
// host's intermediate orders:
//
//  * top levels (subtrees larger than one LDS tile): a per-level,
//    all-subtrees-at-once radix select over the composite key (3 digit
//    launches over the split coordinate, then one launch that resolves
//    the later key words of the few subtrees tied on it, over their tie
//    groups alone), followed by an unordered partition into the next
//    level's compact layout, pivots written straight to their final node
//    slot;
//  * bottom levels: one workgroup per subtree of at most 8191 nodes.  For
//    D <= 4 the rank finish (k_finish_rank) sorts the subtree once per dim
//    in LDS (stable LSD radix passes, ties ordered by the rest of the key)
//    and then only partitions those orders level by level, taking each
//    pivot by rank; explicit split dims and subtrees rooted above level
//    D-1 keep the orders exact by repairing tie runs level by level.
//    Otherwise (D > 4, or long/dense coordinate ties) the bitonic finish
//    (k_finish) puts the subtree in composite-key order and runs every
//    level as a bitonic sort of (coord, position) keys inside power-of-two
//    aligned blocks -- the position tie-break makes each level's sort
//    stable, which reproduces the host's tie chain.  Both write each pivot
//    to its node slot.
//
// All launches are stream-ordered, with no host synchronisation.
#include <stdarg.h>
#include <stdio.h>

#include "cukd_common.h"
#include "block_sort.h"
#include "forest_plan.h"

#include <string.h>
#include <vector>

namespace cukd_amd {

// Partition workgroups: SEL_NT threads x SEL_ITEMS elements per tile (a
// 10M float3 build measured 5.0 ms with 512 x 4, 5.2 ms with 256 x 8 and
// 5.7 ms with 128 x 16: fewer registers per thread, more waves per CU).
enum { SEL_NT = 512, SEL_ITEMS = 4, SEL_BATCH = 8, DIGIT_MAX_BINS = 2048 };
// Select passes: 512 threads on each tile of keys (5.27 ms against 5.56 ms
// with 256 threads and 6.0 ms with 1024, before the partition change).
enum { SELECT_NT = 512 };
// per subtree: DIGIT_MAX_BINS counters, then one bit per bin that some tile
// touched in the current pass
enum { HIST_STRIDE = DIGIT_MAX_BINS + DIGIT_MAX_BINS / 32 };
enum { SEL_HCOPIES = 8 };  // most histogram copies per subtree
enum { MAX_WORDS = CUKD_MAX_DIMS + 1 };

// Per-subtree selection state of one level.
struct SegState {
  uint32_t krem;      // rank of the pivot among the remaining candidates
  uint32_t word;      // word being resolved (or resolved word if resolved)
  uint32_t prefix;    // high bits of 'word' fixed so far
  uint32_t resolved;  // words[0..word] identify the pivot uniquely
  uint32_t tiles_done;
  uint32_t gcount;    // tie-group elements gathered on this level
  uint32_t dim;       // split dim of this subtree
  uint32_t chain;     // split dims, most recent first, 4 bits each
  uint32_t chain_len; // number of coordinate words in the key
  uint32_t fixed[MAX_WORDS];
  uint32_t box[2 * CUKD_MAX_DIMS]; // domain (explicit dims), raw scalar bits
};

__device__ inline uint32_t chain_at(uint32_t chain, int j) { return (chain >> (4 * j)) & 15u; }

// widestDimension on raw scalar bits (box.h:110-125): ties pick the later dim.
__device__ inline int widest_dim_bits(const uint32_t *box, int D, bool is_float)
{
  int best = 0;
  if (is_float) {
    float wb = 0.f;
    for (int d = 0; d < D; d++) {
      const float w = __uint_as_float(box[D + d]) - __uint_as_float(box[d]);
      if (w < wb) continue;
      wb = w;
      best = d;
    }
  } else {
    long long wb = 0;
    for (int d = 0; d < D; d++) {
      const long long w = (long long)(int)box[D + d] - (long long)(int)box[d];
      if (w < wb) continue;
      wb = w;
      best = d;
    }
  }
  return best;
}

// chain of a child: its own dim first, then the parent's chain without it
__device__ inline void child_chain(uint32_t parent_chain, uint32_t parent_len,
                                   uint32_t dim, uint32_t &chain, uint32_t &len)
{
  chain = dim;
  len = 1;
  for (uint32_t j = 0; j < parent_len; j++) {
    const uint32_t d = chain_at(parent_chain, (int)j);
    if (d == dim) continue;
    chain |= d << (4 * len);
    len++;
  }
}

// ---------------------------------------------------------------------------
// Forests: the kernels below build any number of independent trees stored
// back to back in one array in the same launches.  A workgroup finds its
// tree by binary search in a small table; a single tree is passed inline,
// without a table.
// ---------------------------------------------------------------------------

template <class E>
struct TreeSet {
  const E *tab;  // sorted by the searched field; null: 'one' is the only tree
  int count;
  E one;
};

// the entry with the largest key(entry) <= v
template <class E, class K>
__device__ inline E find_tree(const TreeSet<E> &set, int v, K key)
{
  if (!set.tab) return set.one;
  int lo = 0, hi = set.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (key(set.tab[mid]) <= v) lo = mid; else hi = mid - 1;
  }
  return set.tab[lo];
}

template <class E>
__device__ inline E tree_at(const TreeSet<E> &set, int i)
{
  return set.tab ? set.tab[i] : set.one;
}


// ---------------------------------------------------------------------------
// world bounds (builder_common.h:76-163): block min/max of order-preserving
// keys, one atomic per dimension and block.
// ---------------------------------------------------------------------------
__global__ void k_bounds_init(uint32_t *acc, int D, int ntrees)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ntrees * 2 * D) acc[i] = (i % (2 * D)) < D ? 0xffffffffu : 0u;
}

// one workgroup per IN_TILE elements of one tree
__global__ __launch_bounds__(256) void k_bounds(TreeSet<InTree> set, uint32_t *acc,
                                                const uint8_t *points, cukd_layout L)
{
  __shared__ uint32_t s_lo[CUKD_MAX_DIMS][256], s_hi[CUKD_MAX_DIMS][256];
  __shared__ InTree s_tree;
  if (threadIdx.x == 0)
    s_tree = find_tree(set, (int)blockIdx.x, [](const InTree &x) { return x.tile0; });
  __syncthreads();
  const InTree tr = s_tree;
  const int D = L.num_dims;
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const int e0 = ((int)blockIdx.x - tr.tile0) * IN_TILE;
  const int e1 = min(e0 + IN_TILE, tr.n);
  const uint8_t *pts = points + (size_t)tr.base * L.stride;
  uint32_t lo[CUKD_MAX_DIMS], hi[CUKD_MAX_DIMS];
  for (int d = 0; d < D; d++) { lo[d] = 0xffffffffu; hi[d] = 0u; }
  for (int i = e0 + threadIdx.x; i < e1; i += blockDim.x) {
    const uint8_t *e = pts + (size_t)i * L.stride;
    for (int d = 0; d < D; d++) {
      const uint32_t k = ordered_key(load_coord_bits(e, L, d), isf);
      lo[d] = min(lo[d], k);
      hi[d] = max(hi[d], k);
    }
  }
  for (int d = 0; d < D; d++) { s_lo[d][threadIdx.x] = lo[d]; s_hi[d][threadIdx.x] = hi[d]; }
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int d = 0; d < D; d++) {
        s_lo[d][threadIdx.x] = min(s_lo[d][threadIdx.x], s_lo[d][threadIdx.x + off]);
        s_hi[d][threadIdx.x] = max(s_hi[d][threadIdx.x], s_hi[d][threadIdx.x + off]);
      }
    __syncthreads();
  }
  if ((int)threadIdx.x < D) {
    uint32_t *a = acc + (size_t)tr.id * 2 * D;
    atomicMin(&a[threadIdx.x], s_lo[threadIdx.x][0]);
    atomicMax(&a[D + threadIdx.x], s_hi[threadIdx.x][0]);
  }
}

// Decodes every tree's accumulated bounds into the caller's boxes and
// seeds its root state (domain = world box, dim = widest,
// builder_thrust.h:144-156).
__global__ void k_bounds_final(TreeSet<InTree> set, int ntrees, const uint32_t *acc,
                               uint32_t *user_box, SegState *st, cukd_layout L,
                               int explicit_dim)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ntrees) return;
  const InTree tr = tree_at(set, j);
  const int D = L.num_dims;
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  uint32_t box[2 * CUKD_MAX_DIMS];
  for (int i = 0; i < 2 * D; i++) box[i] = unordered_key(acc[(size_t)tr.id * 2 * D + i], isf);
  if (user_box)
    for (int i = 0; i < 2 * D; i++) user_box[(size_t)tr.id * 2 * D + i] = box[i];
  if (st) {
    SegState &root = st[tr.seg];
    for (int i = 0; i < 2 * D; i++) root.box[i] = box[i];
    const int d0 = explicit_dim ? widest_dim_bits(box, D, isf) : 0;
    root.dim = (uint32_t)d0;
    root.chain = (uint32_t)d0;
    root.chain_len = 1;
  }
}

// Counts bin hb (if has) in the LDS histogram h: one atomic per wave when
// every counted lane hits the same bin (the common case deep in the tree,
// where a subtree's keys share their high digits).  Every lane of the wave
// calls it.
__device__ inline void lds_count(uint32_t *h, bool has, uint32_t hb)
{
  const int lane = (int)threadIdx.x & 63;
  const uint64_t act = __ballot(has);
  if (!act) return;
  const int first = __ffsll((unsigned long long)act) - 1;
  const uint32_t fb = __shfl(hb, first);
  const uint64_t same = __ballot(has && hb == fb);
  if (same == act) {
    if (lane == first) atomicAdd(&h[fb], (uint32_t)__popcll(act));
  } else if (has) {
    atomicAdd(&h[hb], 1u);
  }
}

// Adds a tile's LDS digit histogram s_h[0..nbins) to a global digit
// histogram row h with atomics: the non-empty bins, and one bitmap bit per
// bin (32 bins a word, after the DIGIT_MAX_BINS counters).
__device__ inline void flush_hist(uint32_t *h, const uint32_t *s_h, int nbins)
{
  const int lane = (int)threadIdx.x & 63;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
    const uint32_t c = s_h[b];
    if (c) atomicAdd(&h[b], c);
    // bins b..: lanes 0-31 and 32-63 of a wave cover two bitmap words
    const uint64_t m = __ballot(c != 0u);
    const uint32_t part = (lane < 32) ? (uint32_t)m : (uint32_t)(m >> 32);
    if ((lane & 31) == 0 && part) atomicOr(&h[DIGIT_MAX_BINS + (b >> 5)], part);
  }
}

// ---------------------------------------------------------------------------
// copy-in: working buffer A gets the points, their input index inside their
// tree, and the level-0 key.
// ---------------------------------------------------------------------------
// Each tile also clears its share of the select's work area (digit
// histograms, cursors, tie bounds: zwords words from zero), which the
// levels' first launches need zeroed.  (A fill kernel rather than
// hipMemsetAsync, so that a forest build can be captured into a graph.)
__global__ __launch_bounds__(256) void k_copy_in(TreeSet<InTree> set, uint8_t *data,
                                                 uint32_t *idx, uint32_t *key,
                                                 const uint8_t *points, cukd_layout L,
                                                 const SegState *st, int explicit_dim,
                                                 uint32_t *zero, size_t zwords)
{
  __shared__ InTree s_tree;
  if (threadIdx.x == 0)
    s_tree = find_tree(set, (int)blockIdx.x, [](const InTree &x) { return x.tile0; });
  for (size_t z = (size_t)blockIdx.x * blockDim.x + threadIdx.x; z < zwords;
       z += (size_t)gridDim.x * blockDim.x)
    zero[z] = 0u;
  __syncthreads();
  const InTree tr = s_tree;
  const int d0 = explicit_dim ? (int)st[tr.seg].dim : 0;
  const int e0 = ((int)blockIdx.x - tr.tile0) * IN_TILE;
  const int e1 = min(e0 + IN_TILE, tr.n);
  const int W = L.stride >> 2;
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  for (int i = e0 + threadIdx.x; i < e1; i += blockDim.x) {
    const size_t g = (size_t)tr.base + i;
    const uint8_t *src = points + g * L.stride;
    uint32_t *dst = (uint32_t *)(data + g * L.stride);
    for (int w = 0; w < W; w++) dst[w] = ((const uint32_t *)src)[w];
    idx[g] = (uint32_t)i;
    if (key) key[g] = ordered_key(load_coord_bits(src, L, d0), isf);
  }
}

// ---------------------------------------------------------------------------
// global levels
// ---------------------------------------------------------------------------

// The start of subtree s of a level (histogram slot hs = the tree's slot
// plus s): its select state, partition cursors and tie-word bounds.
// Round-robin dims and chains are closed forms of the level; explicit
// dims and domains are written by the level above (or the root's bounds).
__device__ inline void init_state(SegState &t, int level, int s, int n, int D, bool expl,
                                  uint32_t *cursors, uint32_t *tb, int hs)
{
  const int node = (1 << level) - 1 + s;
  t.krem = (uint32_t)subtree_size(2 * node + 1, n);
  t.word = 0;
  t.prefix = 0;
  t.resolved = 0;
  t.tiles_done = 0;
  t.gcount = 0;
  if (!expl) {
    t.dim = (uint32_t)(level % D);
    uint32_t chain = 0, len = 0;
    for (int j = level; j >= 0 && (int)len < D; j--) chain |= (uint32_t)(j % D) << (4 * len++);
    t.chain = chain;
    t.chain_len = len;
  }
  cursors[2 * hs] = 0u;
  cursors[2 * hs + 1] = 0u;
  for (int w = 0; w < MAX_WORDS; w++) {
    tb[(size_t)hs * 2 * MAX_WORDS + w] = 0xffffffffu;
    tb[(size_t)hs * 2 * MAX_WORDS + MAX_WORDS + w] = 0u;
  }
}

// Start of level 0 (later levels start in the tie launch of the level
// above).  One thread per segment of the level.
__global__ void k_level_init(TreeSet<LvlTree> set, int nsegs, SegState *st,
                             uint32_t *cursors, uint32_t *tb, int level, int D, int explicit_dim)
{
  const int sg = blockIdx.x * blockDim.x + threadIdx.x;
  if (sg >= nsegs) return;
  const LvlTree tr = find_tree(set, sg, [](const LvlTree &x) { return x.hseg; });
  const int s = sg - tr.hseg;
  init_state(st[tr.seg + s], level, s, tr.n, D, explicit_dim != 0, cursors, tb, sg);
}

// Value of composite-key word w of an element (word 0 is the cached key).
__device__ inline uint32_t key_word(int w, const SegState &t, uint32_t key0,
                                    const uint8_t *elem, uint32_t idx,
                                    const cukd_layout &L, bool isf)
{
  if (w == 0) return key0;
  if (w < (int)t.chain_len)
    return ordered_key(load_coord_bits(elem, L, (int)chain_at(t.chain, w)), isf);
  return idx;
}

// Tiles of size T over the compact layout of one level.  A tile spans at
// most two subtrees because every subtree on a global level holds at least
// T elements.
__device__ inline int tiles_of(const LevelGeom &g, int s, int T)
{
  const int a = g.offset(s);
  const int b = a + g.size(s) - 1;
  return b / T - a / T + 1;
}

// The bin of h[0..nbins) that holds rank krem, into s_pick as (bin, count
// of lower bins, count of the bin); every thread of the workgroup calls it.
// The threads sum runs of nbins / SELECT_NT consecutive bins into s_part;
// wave 0 alone then scans those sums (SELECT_NT / 64 per lane, then
// shuffles across the lanes), and the owning lane walks down to the bin,
// so the workgroup meets at two barriers only.  nbins is a multiple of
// SELECT_NT.
__device__ inline void pick_bin(const uint32_t *h, int nbins, uint32_t krem, uint32_t *s_part,
                                uint32_t *s_pick)
{
  const int per = nbins / SELECT_NT;
  const int b0 = (int)threadIdx.x * per;
  uint32_t mine = 0;
  for (int i = 0; i < per; i++) mine += h[b0 + i];
  s_part[threadIdx.x] = mine;
  __syncthreads();
  if (threadIdx.x < 64) {
    enum { PL = SELECT_NT / 64 };
    const int lane = (int)threadIdx.x;
    uint32_t lsum = 0;
    for (int i = 0; i < PL; i++) lsum += s_part[lane * PL + i];
    uint32_t incl = lsum;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t v = __shfl_up(incl, off);
      if (lane >= off) incl += v;
    }
    if (krem >= incl - lsum && krem < incl) {
      uint32_t before = incl - lsum;
      int th = lane * PL;
      for (int i = 0; i < PL; i++) {
        th = lane * PL + i;
        const uint32_t c = s_part[th];
        if (before + c > krem) break;
        before += c;
      }
      uint32_t cnt = 0;
      int b = th * per;
      for (int i = 0; i < per; i++) {
        b = th * per + i;
        cnt = h[b];
        if (before + cnt > krem) break;
        before += cnt;
      }
      s_pick[0] = (uint32_t)b;
      s_pick[1] = before;
      s_pick[2] = cnt;
    }
  }
  __syncthreads();
}

// Child domains and split dims of an explicit-dim subtree once its pivot's
// split coordinate (key word 0) is known (updateTagAndSetDim,
// builder_thrust.h:259-314): the pivot coordinate clips the parent box on
// the split dim, and each child splits its widest dim.
__device__ inline void write_child_domains(const SegState &t, SegState *child_st, int s,
                                           const cukd_layout &L, bool isf)
{
  const int D = L.num_dims;
  const uint32_t pc = unordered_key(t.fixed[0], isf);
  for (int side = 0; side < 2; side++) {
    SegState &c = child_st[2 * s + side];
    uint32_t box[2 * CUKD_MAX_DIMS];
    for (int i = 0; i < 2 * D; i++) box[i] = t.box[i];
    box[(side == 0 ? D : 0) + (int)t.dim] = pc;
    const int cd = widest_dim_bits(box, D, isf);
    for (int i = 0; i < 2 * D; i++) c.box[i] = box[i];
    c.dim = (uint32_t)cd;
    child_chain(t.chain, t.chain_len, (uint32_t)cd, c.chain, c.chain_len);
  }
}

// Digit histograms of key word 0: digit d (bits 31-21, 20-10, 9-0) of
// subtree s, copy c, is row (hseg + s) * hcopies + c of buffer d: the bins,
// then one bitmap bit per bin that some tile touched.
__device__ inline const uint32_t *hist_rows(const uint32_t *h, int row0, int hcopies)
{
  return h + (size_t)row0 * hcopies * HIST_STRIDE;
}

// The subtree's digit histogram summed over its copies into h[0..nbins);
// only bins that some tile touched are read.  Every thread calls it.
__device__ inline void load_hist(uint32_t *h, const uint32_t *rows, int hcopies, int nbins)
{
  const int per = nbins / SELECT_NT;  // 2 or 4 consecutive bins, within one bitmap word
  const int b0 = (int)threadIdx.x * per;
  uint32_t sum[DIGIT_MAX_BINS / SELECT_NT];
  for (int i = 0; i < per; i++) sum[i] = 0u;
  for (int c = 0; c < hcopies; c++) {
    const uint32_t *r = rows + (size_t)c * HIST_STRIDE;
    const uint32_t bits = (r[DIGIT_MAX_BINS + (b0 >> 5)] >> (b0 & 31)) & ((1u << per) - 1u);
    for (int i = 0; i < per; i++)
      if ((bits >> i) & 1u) sum[i] += r[b0 + i];
  }
  for (int i = 0; i < per; i++) h[b0 + i] = sum[i];
}

// Word 0's digit pick d of one subtree, replayed from the histogram that
// digit pass d left (its atomics were performed before this launch
// began): krem and prefix advance past digit d.  Returns the count of the
// picked bin.  Every thread calls it and gets the same result; s_h is LDS
// scratch of DIGIT_MAX_BINS words.
__device__ inline uint32_t replay_digit(uint32_t &krem, uint32_t &prefix, const uint32_t *hd,
                                        int row0, int hcopies, int d, uint32_t *s_h,
                                        uint32_t *s_part, uint32_t *s_pick)
{
  const int bits = d == 2 ? 10 : 11;
  const int nbins = 1 << bits;
  load_hist(s_h, hist_rows(hd, row0, hcopies), hcopies, nbins);
  __syncthreads();
  // pick_bin reads s_h and ends with a barrier, so s_h may be reused after
  pick_bin(s_h, nbins, krem, s_part, s_pick);
  krem -= s_pick[1];
  prefix = (prefix << bits) | s_pick[0];
  return s_pick[2];
}

// The state of word 0's select entering digit d of subtree s: for d = 0
// the level's start; after that, what the launch of digit pass d recorded
// in rk when it replayed digit d - 1 (per subtree of the level: krem and
// prefix after digit 0, then after digit 1).  It is written by the tile
// holding the subtree's first element and read only by later launches.
__device__ inline void word0_state(int d, const SegState *st, const uint32_t *rk, int s,
                                   int row, uint32_t &krem, uint32_t &prefix)
{
  if (d == 0) {
    krem = st[s].krem;
    prefix = 0u;
  } else {
    krem = rk[(size_t)row * 4 + 2 * (d - 1)];
    prefix = rk[(size_t)row * 4 + 2 * (d - 1) + 1];
  }
}

// Digit pass d of the radix select on key word 0 (the split coordinate),
// over every subtree of the level.  Every tile first replays the picks of
// the earlier digits from their histograms, then counts digit d of its
// candidates (the elements whose key agrees with those picks) and adds its
// non-empty bins, plus their bitmap words, to the subtree's digit-d
// histogram with atomics.  No tile reads another tile's counts in the same
// launch: the next launch picks, after the kernel boundary has made every
// tile's atomics visible, so the passes need no fences.
//
// Each 16K-key tile loads 8 keys per thread before using any; a wave whose
// counted lanes all hit one bin does a single LDS atomic.  Subtrees with
// many tiles spread their atomics over hcopies histogram copies.
__global__ __launch_bounds__(SELECT_NT) void k_select_digit(
    TreeSet<LvlTree> set, int level, int d, const uint32_t *key, const SegState *st,
    uint32_t *h0, uint32_t *h1, uint32_t *h2, uint32_t *rk, int hcopies)
{
  __shared__ uint32_t s_hist[2][DIGIT_MAX_BINS];
  __shared__ uint32_t s_part[SELECT_NT];
  __shared__ uint32_t s_pick[3];
  __shared__ LvlTree s_tree;
  if (threadIdx.x == 0)
    s_tree = find_tree(set, (int)blockIdx.x, [](const LvlTree &x) { return x.stile0; });
  __syncthreads();
  const LvlTree tr = s_tree;
  LevelGeom g;
  g.init(tr.n, level);
  const int T = tr.tsel;
  key += tr.base;
  st += tr.seg;
  const int t0 = ((int)blockIdx.x - tr.stile0) * T;
  const int t1 = min(t0 + T, g.n_elems);
  const int seg0 = g.seg_of(t0);
  const int seg1 = g.seg_of(t1 - 1);
  const int nslots = seg1 != seg0 ? 2 : 1;
  // the earlier digits' picks: the last one replayed here, and recorded
  // for the next launch by the tile holding the subtree's first element
  uint32_t prefix[2] = {0u, 0u};
  if (d > 0) {
    for (int sl = 0; sl < nslots; sl++) {
      const int s = sl ? seg1 : seg0;
      uint32_t krem;
      word0_state(d - 1, st, rk, s, tr.hseg + s, krem, prefix[sl]);
      replay_digit(krem, prefix[sl], d == 1 ? h0 : h1, tr.hseg + s, hcopies, d - 1,
                   &s_hist[0][0], s_part, s_pick);
      if (threadIdx.x == 0 && (sl == 1 || g.offset(s) >= t0)) {
        rk[(size_t)(tr.hseg + s) * 4 + 2 * (d - 1)] = krem;
        rk[(size_t)(tr.hseg + s) * 4 + 2 * (d - 1) + 1] = prefix[sl];
      }
    }
  }
  const int shift = d == 0 ? 21 : (d == 1 ? 10 : 0);
  const int bits = d == 2 ? 10 : 11;
  const int nbins = 1 << bits;
  const uint32_t mask = (uint32_t)nbins - 1u;
  const int hi_shift = shift + bits;
  for (int i = threadIdx.x; i < 2 * nbins; i += blockDim.x) s_hist[i / nbins][i % nbins] = 0u;
  __syncthreads();
  // tiles inside one subtree use slot 0 only
  const int seg1_begin = nslots == 2 ? g.offset(seg1) : g.n_elems;
  uint32_t *hflat = &s_hist[0][0];
  for (int base = t0; base < t1; base += SELECT_NT * SEL_BATCH) {
    // all of a thread's key loads in flight before any is used
    uint32_t kv[SEL_BATCH];
#pragma unroll
    for (int u = 0; u < SEL_BATCH; u++) {
      const int p = base + u * SELECT_NT + (int)threadIdx.x;
      kv[u] = p < t1 ? key[p] : 0u;
    }
#pragma unroll
    for (int u = 0; u < SEL_BATCH; u++) {
      const int p = base + u * SELECT_NT + (int)threadIdx.x;
      const int slot = p >= seg1_begin ? 1 : 0;
      const uint32_t v = kv[u];
      const bool has = p < t1 && (hi_shift >= 32 || (v >> hi_shift) == prefix[slot]);
      lds_count(hflat, has, (uint32_t)slot * DIGIT_MAX_BINS + ((v >> shift) & mask));
    }
  }
  __syncthreads();
  uint32_t *hd = d == 0 ? h0 : (d == 1 ? h1 : h2);
  const int hcopy = (int)blockIdx.x & (hcopies - 1);
  for (int sl = 0; sl < nslots; sl++) {
    const int s = sl ? seg1 : seg0;
    flush_hist(hd + ((size_t)(tr.hseg + s) * hcopies + hcopy) * HIST_STRIDE, s_hist[sl], nbins);
  }
}

// Pivots of a level: word 0's last pick, then the tied subtrees, in one
// launch.  Every tile replays word 0's three digit picks for its subtrees.
// A subtree whose pivot key is unique is done: the tile that holds its
// first element writes its final state (to st_fin, which no tile of this
// launch reads) and, for explicit dims, its children's domains.  A subtree
// that is not resolved holds several elements whose key word 0 equals its
// pivot's.  Its pivot is the one of rank krem among them in the
// order of the remaining key words: the split-dim chain's earlier
// coordinates, then the input index, which is the order the host builder's
// stable sorts leave (builder_thrust.h:126-142).
//
// Every tile finds its tie-group elements, reserves room for them with one
// atomic per subtree and round, and stores their remaining key words in
// planes (plane j holds word j + 1 of every member).  The planes live in
// the subtree's own range of a buffer that the level leaves unused (the
// next level's element array, which has at least as many words per element
// as there are tie words), so a group of any size fits.  The tile that
// completes the subtree then runs the digit passes over those planes
// alone, inside its workgroup.  Protocol: after the workgroup barrier,
// thread 0 of every tile bumps the subtree's tile counter with an
// agent-scope release RMW; the tile that reads the final count issues an
// agent-scope acquire fence, so (release sequence on the counter, fence
// synchronisation) every tile's appends happen-before its reads.  The
// gather also keeps each tie word's smallest and largest value over the
// group (tb, per subtree of the level): a word on which the whole group
// agrees is fixed without a pass, and so is every leading digit that its
// bounds share (all-equal points thus cost the index word's low digits
// only).  With no tied subtree on the level the launch returns at once.
enum { TIE_BATCH = 8 };
__global__ __launch_bounds__(SELECT_NT) void k_tie_resolve(
    TreeSet<LvlTree> set, int level, const uint32_t *key, const uint8_t *data,
    const uint32_t *idx, SegState *st, SegState *st_fin, const uint32_t *h2, const uint32_t *rk,
    int hcopies, uint8_t *scratch, uint32_t *tb, uint32_t *cursors_next, uint32_t *tb_next,
    cukd_layout L, int explicit_dim)
{
  enum { NW = SELECT_NT / 64 };
  __shared__ uint32_t s_hist[DIGIT_MAX_BINS];
  __shared__ SegState s_st[2];
  __shared__ int s_last[2];
  __shared__ uint32_t s_part[SELECT_NT];
  __shared__ uint32_t s_pick[3];
  __shared__ uint32_t s_wcnt[2][NW];  // members of one round per (slot, wave), then their offsets
  __shared__ uint32_t s_tmin[2][MAX_WORDS], s_tmax[2][MAX_WORDS];  // tie-word bounds of the tile
  __shared__ LvlTree s_tree;
  if (threadIdx.x == 0)
    s_tree = find_tree(set, (int)blockIdx.x, [](const LvlTree &x) { return x.stile0; });
  __syncthreads();
  const LvlTree tr = s_tree;
  LevelGeom g;
  g.init(tr.n, level);
  const int T = tr.tsel;
  const int W = L.stride >> 2;
  key += tr.base;
  data += (size_t)tr.base * L.stride;
  idx += tr.base;
  uint32_t *words = (uint32_t *)(scratch + (size_t)tr.base * L.stride);
  SegState *child_st = st + tr.cseg;
  st += tr.seg;
  st_fin += tr.hseg;
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const int t0 = ((int)blockIdx.x - tr.stile0) * T;
  const int t1 = min(t0 + T, g.n_elems);
  const int seg0 = g.seg_of(t0);
  const int seg1 = g.seg_of(t1 - 1);
  if (threadIdx.x == 0) s_st[0] = st[seg0];
  if (threadIdx.x == 1) s_st[1] = st[seg1];
  if (threadIdx.x < 2 * MAX_WORDS) {
    (&s_tmin[0][0])[threadIdx.x] = 0xffffffffu;
    (&s_tmax[0][0])[threadIdx.x] = 0u;
  }
  __syncthreads();
  // tiles inside one subtree use slot 0 only
  const int nslots = seg1 != seg0 ? 2 : 1;
  // word 0's picks: the pivot's key word 0, and its rank among the
  // elements that share it
  for (int sl = 0; sl < nslots; sl++) {
    const int s = sl ? seg1 : seg0;
    uint32_t krem, key0;
    word0_state(2, st, rk, s, tr.hseg + s, krem, key0);
    const uint32_t cnt =
        replay_digit(krem, key0, h2, tr.hseg + s, hcopies, 2, s_hist, s_part, s_pick);
    if (threadIdx.x == 0) {
      SegState &t = s_st[sl];
      t.krem = krem;
      t.fixed[0] = key0;
      t.prefix = 0u;
      t.resolved = cnt == 1u ? 1u : 0u;
      t.word = cnt == 1u ? 0u : 1u;
    }
    __syncthreads();
  }
  // the tile holding a subtree's first element writes what depends on key
  // word 0 alone
  if (threadIdx.x < 2 && (int)threadIdx.x < nslots) {
    const int sl = (int)threadIdx.x;
    const int s = sl ? seg1 : seg0;
    if (g.offset(s) >= t0) {
      if (explicit_dim) write_child_domains(s_st[sl], child_st, s, L, isf);
      if (s_st[sl].resolved) st_fin[s] = s_st[sl];
      // the children's start on the next top level (their slots there
      // follow the tree's next-level slot)
      if (tr.more)
        for (int side = 0; side < 2; side++)
          init_state(child_st[2 * s + side], level + 1, 2 * s + side, tr.n, L.num_dims,
                     explicit_dim != 0, cursors_next, tb_next, tr.nhseg + 2 * s + side);
    }
  }
  const bool tied0 = !s_st[0].resolved;
  const bool tied1 = nslots == 2 && !s_st[1].resolved;
  if (!tied0 && !tied1) return;
  const int seg1_begin = nslots == 2 ? g.offset(seg1) : g.n_elems;
  const int twmax = max((int)s_st[0].chain_len, nslots == 2 ? (int)s_st[1].chain_len : 0);
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  const uint64_t below = (1ull << lane) - 1ull;
  // gather
  for (int base = t0; base < t1; base += SELECT_NT * TIE_BATCH) {
    uint32_t kv[TIE_BATCH];
#pragma unroll
    for (int u = 0; u < TIE_BATCH; u++) {
      const int p = base + u * SELECT_NT + (int)threadIdx.x;
      const int slot = p >= seg1_begin ? 1 : 0;
      kv[u] = p < t1 && (slot ? tied1 : tied0) ? key[p] : 0u;
    }
    bool in[TIE_BATCH];
    uint32_t rank[TIE_BATCH];
    uint32_t wcnt[2] = {0u, 0u};
#pragma unroll
    for (int u = 0; u < TIE_BATCH; u++) {
      const int p = base + u * SELECT_NT + (int)threadIdx.x;
      const int slot = p >= seg1_begin ? 1 : 0;
      in[u] = p < t1 && (slot ? tied1 : tied0) && kv[u] == s_st[slot].fixed[0];
      rank[u] = 0u;
      for (int sl = 0; sl < nslots; sl++) {
        const uint64_t m = __ballot(in[u] && slot == sl);
        if (in[u] && slot == sl) rank[u] = wcnt[sl] + (uint32_t)__popcll(m & below);
        wcnt[sl] += (uint32_t)__popcll(m);
      }
    }
    if (lane == 0) {
      s_wcnt[0][wave] = wcnt[0];
      s_wcnt[1][wave] = wcnt[1];
    }
    __syncthreads();
    if (threadIdx.x < 2 && (int)threadIdx.x < nslots) {
      // one reservation per subtree and round; the waves' offsets follow
      const int sl = (int)threadIdx.x;
      uint32_t tot = 0u;
      for (int w = 0; w < NW; w++) {
        const uint32_t c = s_wcnt[sl][w];
        s_wcnt[sl][w] = tot;
        tot += c;
      }
      const uint32_t b = tot ? atomicAdd(&st[sl ? seg1 : seg0].gcount, tot) : 0u;
      for (int w = 0; w < NW; w++) s_wcnt[sl][w] += b;
    }
    __syncthreads();
    // the members' tie words into the planes and the tile's word bounds
    // (one LDS update per wave where the wave's members agree)
#pragma unroll
    for (int u = 0; u < TIE_BATCH; u++) {
      const int p = base + u * SELECT_NT + (int)threadIdx.x;
      const int slot = p >= seg1_begin ? 1 : 0;
      const SegState &t = s_st[slot];
      const int s = slot ? seg1 : seg0;
      const uint32_t j = in[u] ? s_wcnt[slot][wave] + rank[u] : 0u;
      const int size = g.size(s);
      uint32_t *reg = words + (size_t)g.offset(s) * W;
      const uint8_t *elem = data + (size_t)(in[u] ? p : t0) * L.stride;
      const uint32_t id = in[u] ? idx[p] : 0u;
      for (int w = 1; w <= twmax; w++) {
        const bool mine = in[u] && w <= (int)t.chain_len;
        const uint32_t v = mine ? key_word(w, t, kv[u], elem, id, L, isf) : 0u;
        if (mine) reg[(size_t)(w - 1) * size + j] = v;
        for (int sl = 0; sl < nslots; sl++) {
          const bool c = mine && slot == sl;
          const uint64_t act = __ballot(c);
          if (!act) continue;
          const int first = __ffsll((unsigned long long)act) - 1;
          const uint32_t fv = __shfl(v, first);
          if (__ballot(c && v == fv) == act) {
            if (lane == first) {
              atomicMin(&s_tmin[sl][w], fv);
              atomicMax(&s_tmax[sl][w], fv);
            }
          } else if (c) {
            atomicMin(&s_tmin[sl][w], v);
            atomicMax(&s_tmax[sl][w], v);
          }
        }
      }
    }
    __syncthreads();
  }
  // the tile's word bounds into the subtree's
  if ((int)threadIdx.x < nslots * MAX_WORDS) {
    const int sl = (int)threadIdx.x / MAX_WORDS, w = (int)threadIdx.x % MAX_WORDS;
    uint32_t *b = tb + (size_t)(tr.hseg + (sl ? seg1 : seg0)) * 2 * MAX_WORDS;
    if (s_tmin[sl][w] <= s_tmax[sl][w]) {
      atomicMin(&b[w], s_tmin[sl][w]);
      atomicMax(&b[MAX_WORDS + w], s_tmax[sl][w]);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int sl = 0; sl < 2; sl++) {
      s_last[sl] = 0;
      if (sl < nslots && (sl ? tied1 : tied0)) {
        const int s = sl ? seg1 : seg0;
        // release: this tile's planes happen-before the last tile's reads
        const uint32_t done = __hip_atomic_fetch_add(&st[s].tiles_done, 1u, __ATOMIC_RELEASE,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        s_last[sl] = (int)done == tiles_of(g, s, T) - 1;
      }
    }
  }
  __syncthreads();
  for (int sl = 0; sl < nslots; sl++) {
    if (!s_last[sl]) continue;
    // every thread acquires before reading the other tiles' planes
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int s = sl ? seg1 : seg0;
    const uint32_t m = __hip_atomic_load(&st[s].gcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const size_t size = (size_t)g.size(s);
    const uint32_t *reg = words + (size_t)g.offset(s) * W;
    const uint32_t *bnd = tb + (size_t)(tr.hseg + s) * 2 * MAX_WORDS;
    SegState t = s_st[sl];
    for (int w = (int)t.word; w <= (int)t.chain_len && !t.resolved; w++) {
      const uint32_t *plane = reg + (size_t)(w - 1) * size;
      const uint32_t lo = __hip_atomic_load(&bnd[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t hi =
          __hip_atomic_load(&bnd[MAX_WORDS + w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t prefix = 0u, cnt = 0u;
      if (lo == hi) {
        // the whole group agrees on this word: still tied (cnt >= 2)
        t.fixed[w] = lo;
        t.word = (uint32_t)w;
        continue;
      }
      for (int d = 0; d < 3; d++) {
        const int shift = d == 0 ? 21 : (d == 1 ? 10 : 0);
        const int bits = d == 2 ? 10 : 11;
        const int nbins = 1 << bits;
        const uint32_t mask = (uint32_t)nbins - 1u;
        const int hi_shift = shift + bits;
        if ((lo >> shift) == (hi >> shift)) {
          // every candidate is in one bin: krem and the candidates stay
          // (never the last digit, since lo != hi)
          prefix = (hi_shift < 32 ? (prefix << bits) : 0u) | ((lo >> shift) & mask);
          continue;
        }
        for (int i = threadIdx.x; i < nbins; i += blockDim.x) s_hist[i] = 0u;
        __syncthreads();
        for (uint32_t i0 = 0; i0 < m; i0 += SELECT_NT * TIE_BATCH) {
          uint32_t v[TIE_BATCH];
#pragma unroll
          for (int u = 0; u < TIE_BATCH; u++) {
            const uint32_t i = i0 + u * SELECT_NT + threadIdx.x;
            v[u] = i < m ? plane[i] : 0u;
          }
#pragma unroll
          for (int u = 0; u < TIE_BATCH; u++) {
            const uint32_t i = i0 + u * SELECT_NT + threadIdx.x;
            bool has = i < m;
            // candidates agree with the pivot on the earlier tie words
            for (int w2 = 1; w2 < w && has; w2++)
              has = reg[(size_t)(w2 - 1) * size + i] == t.fixed[w2];
            if (has && hi_shift < 32 && (v[u] >> hi_shift) != prefix) has = false;
            const uint32_t hb = (v[u] >> shift) & mask;
            const uint64_t act = __ballot(has);
            if (act) {
              const int first = __ffsll((unsigned long long)act) - 1;
              const uint32_t fb = __shfl(hb, first);
              const uint64_t same = __ballot(has && hb == fb);
              if (same == act) {
                if (lane == first) atomicAdd(&s_hist[fb], (uint32_t)__popcll(act));
              } else if (has) {
                atomicAdd(&s_hist[hb], 1u);
              }
            }
          }
        }
        __syncthreads();
        pick_bin(s_hist, nbins, t.krem, s_part, s_pick);
        t.krem -= s_pick[1];
        cnt = s_pick[2];
        prefix = (hi_shift < 32 ? (prefix << bits) : 0u) | s_pick[0];
      }
      t.fixed[w] = prefix;
      t.word = (uint32_t)w;
      if (cnt == 1u) t.resolved = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      t.tiles_done = 0;
      t.gcount = 0;
      t.prefix = 0;
      st_fin[s] = t;
    }
    __syncthreads();
  }
}

// Partition of every subtree around its pivot: pivots go to their node
// slot in the output, the rest to the two child subtrees of the next
// level's compact layout (order inside a child is irrelevant: the next
// level selects by explicit key).
//
// SW > 0 (elements of at most SW words, SW = 3 or 4): the tile's elements
// and indices are staged in LDS and written out in destination order, so
// consecutive lanes store consecutive words of one child run; in source
// order (SW = 0, wider elements), the stores of a tile scatter over four
// runs and the L2s take one request per few words.  The staging buffer is
// sized for SW words so that 3-word points keep four workgroups per CU.
template <int SW>
__global__ __launch_bounds__(SEL_NT) void k_partition(
    TreeSet<LvlTree> set, int level, const uint32_t *key, const uint8_t *data,
    const uint32_t *idx, const SegState *st, const SegState *st_fin, uint8_t *ndata,
    uint32_t *nidx, uint32_t *nkey, uint32_t *cursors, uint8_t *out, cukd_layout L,
    int explicit_dim, uint32_t *h0, uint32_t *h1, uint32_t *h2, int hrows)
{
  enum { TILE = SEL_ITEMS * SEL_NT, STAGED = SW > 0 };
  __shared__ SegState s_st[2];
  __shared__ uint32_t s_cdim[4];
  __shared__ uint32_t s_cnt[SEL_ITEMS][SEL_NT / 64][4];
  __shared__ uint32_t s_base[4];
  __shared__ uint32_t s_tot[4];  // elements of each class in the tile
  // source order: per tile element, child slot or -(node+1) for the pivot
  __shared__ int s_dst[STAGED ? 1 : TILE];
  __shared__ uint8_t s_edim[STAGED ? 1 : TILE];  // child key dim, or the pivot's split dim
  // staged: the tile's words and indices, and the element at each
  // destination-order position
  __shared__ uint32_t s_tile[STAGED ? TILE * SW : 1];
  __shared__ uint32_t s_idx[STAGED ? TILE : 1];
  __shared__ uint16_t s_inv[STAGED ? TILE : 1];
  __shared__ int s_pe[2];  // staged: tile element of each slot's pivot, or -1
  __shared__ LvlTree s_tree;
  if (threadIdx.x == 0)
    s_tree = find_tree(set, (int)blockIdx.x, [](const LvlTree &x) { return x.ptile0; });
  __syncthreads();
  const LvlTree tr = s_tree;
  LevelGeom g, gn;
  g.init(tr.n, level);
  gn.init(tr.n, level + 1);
  const int T = tr.tpart;
  const int write_next_key = tr.more;
  key += tr.base;
  data += (size_t)tr.base * L.stride;
  idx += tr.base;
  ndata += (size_t)tr.base * L.stride;
  nidx += tr.base;
  nkey += tr.base;
  out += (size_t)tr.base * L.stride;
  cursors += 2 * tr.hseg;
  const SegState *child_st = st + tr.cseg;
  st_fin += tr.hseg;
  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const int D = L.num_dims;
  const int t0 = ((int)blockIdx.x - tr.ptile0) * T;
  const int t1 = min(t0 + T, g.n_elems);
  const int seg0 = g.seg_of(t0);
  const int seg1 = g.seg_of(t1 - 1);
  if (threadIdx.x == 0) s_st[0] = st_fin[seg0];
  if (threadIdx.x == 1) s_st[1] = st_fin[seg1];
  // this level's digit histograms are read for the last time by now: clear
  // the touched bins (bitmaps) for the next level, a few rows per tile
  for (int r = (int)blockIdx.x; r < hrows; r += (int)gridDim.x) {
    if (threadIdx.x < 3 * (DIGIT_MAX_BINS / 32)) {
      const int hsel = (int)threadIdx.x / (DIGIT_MAX_BINS / 32);
      const int w = (int)threadIdx.x % (DIGIT_MAX_BINS / 32);
      uint32_t *row = (hsel == 0 ? h0 : (hsel == 1 ? h1 : h2)) + (size_t)r * HIST_STRIDE;
      uint32_t bm = row[DIGIT_MAX_BINS + w];
      if (bm) {
        row[DIGIT_MAX_BINS + w] = 0u;
        while (bm) {
          row[w * 32 + __builtin_ctz(bm)] = 0u;
          bm &= bm - 1u;
        }
      }
    }
  }
  if (threadIdx.x < 4) {
    const int sl = threadIdx.x >> 1, side = threadIdx.x & 1;
    const int s = sl ? seg1 : seg0;
    s_cdim[threadIdx.x] = explicit_dim ? child_st[2 * s + side].dim
                                       : (uint32_t)((g.L + 1) % D);
  }
  __syncthreads();
  // tiles inside one subtree use slot 0 only
  const int seg1_begin = seg1 != seg0 ? g.offset(seg1) : g.n_elems;
  const int W = L.stride >> 2;
  const int nel = t1 - t0;
  if constexpr (STAGED) {
    // the tile's words and indices into LDS (loads issued before the
    // classification's, all in flight together)
    const uint32_t *src = (const uint32_t *)(data + (size_t)t0 * L.stride);
    const int nw = nel * W;
    uint32_t v[SW * SEL_ITEMS];
#pragma unroll
    for (int u = 0; u < SW * SEL_ITEMS; u++) {
      const int i = (int)threadIdx.x + u * SEL_NT;
      v[u] = i < nw ? src[i] : 0u;
    }
    uint32_t vi[SEL_ITEMS];
#pragma unroll
    for (int u = 0; u < SEL_ITEMS; u++) {
      const int e = (int)threadIdx.x + u * SEL_NT;
      vi[u] = e < nel ? idx[t0 + e] : 0u;
    }
#pragma unroll
    for (int u = 0; u < SW * SEL_ITEMS; u++) {
      const int i = (int)threadIdx.x + u * SEL_NT;
      if (i < nw) s_tile[i] = v[u];
    }
#pragma unroll
    for (int u = 0; u < SEL_ITEMS; u++) {
      const int e = (int)threadIdx.x + u * SEL_NT;
      if (e < nel) s_idx[e] = vi[u];
    }
    if (threadIdx.x < 2) s_pe[threadIdx.x] = -1;
  }
  // classify (all key loads in flight first; the element and index only
  // for subtrees still tied on a later word)
  uint32_t kv[SEL_ITEMS];
#pragma unroll
  for (int k = 0; k < SEL_ITEMS; k++) {
    const int p = t0 + k * SEL_NT + threadIdx.x;
    kv[k] = p < t1 ? key[p] : 0u;
  }
  uint8_t cls[SEL_ITEMS]; // bit0-1: side (0 left, 1 right, 2 pivot), bit2: slot
#pragma unroll
  for (int k = 0; k < SEL_ITEMS; k++) {
    const int p = t0 + k * SEL_NT + threadIdx.x;
    cls[k] = 0xff;
    if (p >= t1) continue;
    const int slot = p >= seg1_begin ? 1 : 0;
    const SegState &t = s_st[slot];
    const uint32_t k0 = kv[k];
    int side = k0 == t.fixed[0] ? 2 : (k0 < t.fixed[0] ? 0 : 1);
    if (side == 2 && t.word > 0) {
      const uint8_t *elem = data + (size_t)p * L.stride;
      const uint32_t id = idx[p];
      for (int w = 1; w <= (int)t.word; w++) {
        const uint32_t v = key_word(w, t, k0, elem, id, L, isf);
        if (v != t.fixed[w]) { side = v < t.fixed[w] ? 0 : 1; break; }
      }
    }
    cls[k] = (uint8_t)(side | (slot << 2));
  }
  // Ranks inside each (slot, side) class in (item, wave, lane) order: for
  // one item the lanes of a wave that go the same way write consecutive
  // slots, so every store instruction is contiguous.  (Order inside a
  // child is irrelevant: the next level selects by explicit key.)
  const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
  const uint64_t below = (1ull << lane) - 1ull;
  uint16_t wrank[SEL_ITEMS];
  for (int k = 0; k < SEL_ITEMS; k++) {
    const int c = (cls[k] == 0xff || (cls[k] & 3) == 2) ? -1 : ((cls[k] >> 2) * 2 + (cls[k] & 3));
    wrank[k] = 0;
    for (int cc = 0; cc < 4; cc++) {
      const uint64_t m = __ballot(c == cc);
      if (c == cc) wrank[k] = (uint16_t)__popcll(m & below);
      if (lane == 0) s_cnt[k][wave][cc] = (uint32_t)__popcll(m);
    }
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int cc = (int)threadIdx.x;
    uint32_t run = 0;
    for (int k = 0; k < SEL_ITEMS; k++)
      for (int w = 0; w < SEL_NT / 64; w++) {
        const uint32_t v = s_cnt[k][w][cc];
        s_cnt[k][w][cc] = run;
        run += v;
      }
    const int s = (cc >> 1) ? seg1 : seg0;
    s_base[cc] = run ? atomicAdd(&cursors[2 * s + (cc & 1)], run) : 0u;
    s_tot[cc] = run;
  }
  __syncthreads();
  if constexpr (STAGED) {
    // destination order: class c (slot*2 + side) holds positions
    // [pre[c], pre[c+1]); position pos of class c goes to child slot
    // cbase[c] + pos - pre[c]
    int pre[5];
    int cbase[4];
    pre[0] = 0;
    for (int c = 0; c < 4; c++) {
      pre[c + 1] = pre[c] + (int)s_tot[c];
      const int sg = (c >> 1) ? seg1 : seg0;
      cbase[c] = gn.offset(2 * sg + (c & 1)) + (int)s_base[c];
    }
    for (int k = 0; k < SEL_ITEMS; k++) {
      const int e = k * SEL_NT + (int)threadIdx.x;
      if (cls[k] == 0xff) continue;
      const int side = cls[k] & 3, slot = cls[k] >> 2;
      if (side == 2) {
        s_pe[slot] = e;
        continue;
      }
      const int c = slot * 2 + side;
      s_inv[pre[c] + (int)(s_cnt[k][wave][c] + wrank[k])] = (uint16_t)e;
    }
    __syncthreads();
    const int npos = pre[4];
    const int nw = npos * W;
    const uint32_t magic = (uint32_t)((0x100000000ull + (unsigned)W - 1u) / (unsigned)W);  // pos = t / W
    const int cw0 = L.point_offset >> 2;
    uint32_t *nd = (uint32_t *)ndata;
    for (int t = (int)threadIdx.x; t < nw; t += SEL_NT) {
      const int pos = W == 1 ? t : (int)__umulhi((uint32_t)t, magic);
      const int w = t - pos * W;
      const int c = (pos >= pre[1]) + (pos >= pre[2]) + (pos >= pre[3]);
      const int dst = cbase[c] + pos - pre[c];
      const uint32_t v = s_tile[(int)s_inv[pos] * W + w];
      nd[(size_t)dst * W + w] = v;
      if (write_next_key && w == cw0 + (int)s_cdim[c]) nkey[dst] = ordered_key(v, isf);
    }
    for (int pos = (int)threadIdx.x; pos < npos; pos += SEL_NT) {
      const int c = (pos >= pre[1]) + (pos >= pre[2]) + (pos >= pre[3]);
      nidx[cbase[c] + pos - pre[c]] = s_idx[s_inv[pos]];
    }
    // the pivots to their node slots
    for (int i = (int)threadIdx.x; i < 2 * W; i += SEL_NT) {
      const int sl = i / W, w = i - sl * W;
      const int e = s_pe[sl];
      if (e < 0) continue;
      const int node = (1 << g.L) - 1 + (sl ? seg1 : seg0);
      ((uint32_t *)(out + (size_t)node * L.stride))[w] =
          patch_dim_word(s_tile[e * W + w], w, L, (int)s_st[sl].dim);
    }
    return;
  }
  // destinations, per tile element
  for (int k = 0; k < SEL_ITEMS; k++) {
    const int e = k * SEL_NT + (int)threadIdx.x;
    if (cls[k] == 0xff) continue;
    const int side = cls[k] & 3, slot = cls[k] >> 2;
    const int s = slot ? seg1 : seg0;
    if (side == 2) {
      s_dst[e] = -((1 << g.L) - 1 + s) - 1;
      s_edim[e] = (uint8_t)s_st[slot].dim;
      continue;
    }
    const int c = slot * 2 + side;
    const uint32_t rank = s_cnt[k][wave][c] + wrank[k];
    s_dst[e] = gn.offset(2 * s + side) + (int)(s_base[c] + rank);
    s_edim[e] = (uint8_t)s_cdim[c];
  }
  __syncthreads();
  // Copy word by word: the tile's elements are contiguous in the source,
  // so the loads are coalesced and independent; each element's words go to
  // its destination (child slot, or the pivot's node).
  const int nw = nel * W;
  const uint32_t *src = (const uint32_t *)(data + (size_t)t0 * L.stride);
  const uint32_t magic = (uint32_t)((0x100000000ull + (unsigned)W - 1u) / (unsigned)W);  // e = i / W
  const int cw0 = L.point_offset >> 2;
  enum { CU = 4 };
  for (int i0 = (int)threadIdx.x; i0 < nw; i0 += SEL_NT * CU) {
    uint32_t v[CU];
#pragma unroll
    for (int u = 0; u < CU; u++) {
      const int i = i0 + u * SEL_NT;
      v[u] = i < nw ? src[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < CU; u++) {
      const int i = i0 + u * SEL_NT;
      if (i >= nw) continue;
      const int e = W == 1 ? i : (int)__umulhi((uint32_t)i, magic);
      const int w = i - e * W;
      const int dst = s_dst[e];
      if (dst >= 0) {
        ((uint32_t *)(ndata + (size_t)dst * L.stride))[w] = v[u];
        if (write_next_key && w == cw0 + (int)s_edim[e]) nkey[dst] = ordered_key(v[u], isf);
      } else {
        ((uint32_t *)(out + (size_t)(-dst - 1) * L.stride))[w] = patch_dim_word(v[u], w, L, (int)s_edim[e]);
      }
    }
  }
  for (int e = (int)threadIdx.x; e < nel; e += SEL_NT) {
    const int dst = s_dst[e];
    if (dst >= 0) nidx[dst] = idx[t0 + e];
  }
}

// ---------------------------------------------------------------------------
// LDS finish: one workgroup builds a whole subtree of at most CAP-1 nodes.
// ---------------------------------------------------------------------------

// only: if set, the workgroups whose word is zero have nothing to do (their
// subtrees were finished by the rank finish)
template <int CAP, int NT, bool EXPLICIT>
__global__ __launch_bounds__(NT) void k_finish(
    TreeSet<FinTree> set, const uint8_t *dataA, const uint8_t *dataB, const uint32_t *idxA,
    const uint32_t *idxB, const SegState *st, uint8_t *out, cukd_layout L, const uint32_t *only)
{
  enum { E = CAP / NT };
  if (only && only[blockIdx.x] == 0u) return;
  __shared__ uint32_t s_xv[CAP];  // sort exchange, then the slot scatter
  __shared__ uint32_t s_split[EXPLICIT ? CAP : 1];
  __shared__ uint8_t s_sdim[EXPLICIT ? CAP : 1];
  __shared__ uint8_t s_bdim[EXPLICIT ? CAP / 2 : 1];
  __shared__ uint32_t s_box[2 * CUKD_MAX_DIMS];
  __shared__ FinTree s_tree;
  if (threadIdx.x == 0)
    s_tree = find_tree(set, (int)blockIdx.x, [](const FinTree &x) { return x.blk0; });
  __syncthreads();
  const FinTree tr = s_tree;
  const int n = tr.n, L0 = tr.L0;
  // after L0 partitions a tree's working copy is in buffer A (L0 even) or B
  const uint8_t *data = ((L0 & 1) ? dataB : dataA) + (size_t)tr.base * L.stride;
  const uint32_t *idx = ((L0 & 1) ? idxB : idxA) + tr.base;
  out += (size_t)tr.base * L.stride;
  st += tr.seg;

  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const int D = L.num_dims;
  const int h = num_levels_for(n);
  const int hs = h - L0;            // levels in this subtree
  const int Cs = 1 << hs;           // aligned capacity (<= CAP)
  const int s = (int)blockIdx.x - tr.blk0;
  LevelGeom g;
  g.init(n, L0);
  const int g_node = (1 << L0) - 1 + s;
  const int m = g.size(s);
  const int base = L0 > 0 ? g.offset(s) : 0;
  const uint8_t *sdata = data + (size_t)base * L.stride;
  const uint32_t *sidx = idx + base;
  const SegState &gst = st[s];
  if (EXPLICIT && threadIdx.x < 2 * D) s_box[threadIdx.x] = gst.box[threadIdx.x];

  const int p0 = threadIdx.x * E;
  uint64_t key[E];
  uint32_t slot[E];
#pragma unroll
  for (int e = 0; e < E; e++) slot[e] = (uint32_t)(p0 + e);
  __syncthreads();  // s_box

  // Entry order: the host builder's order inside this subtree before its
  // level-L0 sort is the composite key of the levels above: the subtree's
  // own chain without its first dim, then the input index.  LSD passes:
  // (last chain coordinate, input index) first, then the chain dims from
  // least to most significant, each tie-broken by the current position.
  if (L0 > 0) {
    uint32_t chain = 0, len = 0;
    if (EXPLICIT) {
      chain = gst.chain;
      len = gst.chain_len;
    } else {
      for (int l = L0; l >= 0 && (int)len < D; l--) chain |= (uint32_t)(l % D) << (4 * len++);
    }
    for (int j = (int)len - 1; j >= 0; j--) {
      const bool first = j == (int)len - 1;
      if (j == 0 && !first) break;  // chain[0] is this level's own dim
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int p = p0 + e;
        uint64_t kk = ~0ull;
        if (p < m) {
          const uint32_t sl = slot[e];
          if (first) {
            const uint64_t hi = j >= 1 ? (uint64_t)ordered_key(load_coord_bits(sdata + (size_t)sl * L.stride, L, (int)chain_at(chain, j)), isf) : 0ull;
            kk = (hi << 32) | (uint64_t)sidx[sl];
          } else {
            kk = ((uint64_t)ordered_key(load_coord_bits(sdata + (size_t)sl * L.stride, L, (int)chain_at(chain, j)), isf) << 32) | (uint64_t)p;
          }
        }
        key[e] = kk;
      }
      BlockSort<CAP, NT>::sort(key, slot, Cs, s_xv);
      if (j == 0) break;
    }
  }

  for (int j = 0; j < hs; j++) {
    const int B = Cs >> j;
    const int nblk = 1 << j;
    const int first_node = (g_node + 1) * nblk - 1;
    if (EXPLICIT) {
      for (int b = threadIdx.x; b < nblk; b += NT) {
        if (first_node + b >= n) { s_bdim[b] = 0; continue; }  // no such node
        uint32_t box[2 * CUKD_MAX_DIMS];
        for (int i = 0; i < 2 * D; i++) box[i] = s_box[i];
        int local = nblk - 1 + b;
        while (local > 0) {
          const int parent = (local - 1) >> 1;
          const int pd = s_sdim[parent];
          const uint32_t sv = s_split[parent];
          if (local & 1) {
            const uint32_t u = box[D + pd];
            const bool lt = isf ? (__uint_as_float(u) < __uint_as_float(sv)) : ((int)u < (int)sv);
            box[D + pd] = lt ? u : sv;
          } else {
            const uint32_t l = box[pd];
            const bool lt = isf ? (__uint_as_float(sv) < __uint_as_float(l)) : ((int)sv < (int)l);
            box[pd] = lt ? l : sv;
          }
          local = parent;
        }
        s_bdim[b] = (uint8_t)widest_dim_bits(box, D, isf);
      }
      __syncthreads();
    }
    const int rr_dim = (L0 + j) % D;
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int p = p0 + e;
      uint64_t kk = ~0ull;
      if (p < Cs) {
        const int b = p / B;
        const int sz = subtree_size(first_node + b, n);
        if (p - b * B < sz) {
          const int d = EXPLICIT ? (int)s_bdim[b] : rr_dim;
          kk = ((uint64_t)ordered_key(load_coord_bits(sdata + (size_t)slot[e] * L.stride, L, d), isf) << 32) | (uint64_t)p;
        }
      }
      key[e] = kk;
    }
    if (B > 1) BlockSort<CAP, NT>::sort(key, slot, B, s_xv);
#pragma unroll
    for (int e = 0; e < E; e++) {
      if (key[e] == ~0ull) continue;
      const int q = p0 + e;
      const int b = q / B;
      const int node = first_node + b;
      const int r = q - b * B;
      const int left = subtree_size(2 * node + 1, n);
      const uint32_t sl = slot[e];
      if (r == left) {
        const int d = EXPLICIT ? (int)s_bdim[b] : rr_dim;
        copy_elem(out + (size_t)node * L.stride, sdata + (size_t)sl * L.stride, L, d);
        if (EXPLICIT) {
          s_split[nblk - 1 + b] = load_coord_bits(sdata + (size_t)sl * L.stride, L, d);
          s_sdim[nblk - 1 + b] = (uint8_t)d;
        }
      } else {
        const int np = r < left ? b * B + r : b * B + (B >> 1) + (r - left - 1);
        s_xv[np] = sl;
      }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; e++) slot[e] = s_xv[p0 + e];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// LDS finish by ranks: round-robin dims, D <= RK_DMAX, subtree roots at
// level L0 >= D-1.
// ---------------------------------------------------------------------------
// What k_finish sorts at level l, inside each subtree, is the order of
// the lexicographic key (c[l%D], c[(l-1)%D], ..., c[(l-D+1)%D], input
// index).  With L0 >= D-1 the chain of levels above always covers all D
// dims, and a repeated dim adds nothing, so that key depends on d = l%D
// only: there are D fixed orders key_0 .. key_{D-1}.  This kernel builds
// them once by stable LSD radix passes:
//   key_{D-1} = sort by idx, then stably by c[0], c[1], ..., c[D-1]
//   key_d     = key_{d-1} stably sorted by c[d]   (d = 0 .. D-2)
// (a stable sort by c[d] of (c[d-1], ..., c[d-D]=c[d], idx) order gives
// (c[d], c[d-1], ..., c[d-D+1], idx)).  Each level then takes the pivot of
// every subtree from the order of its dim and stably partitions the other
// orders by child, so the orders stay sorted inside every subtree.  Nodes,
// pivots and tie-breaks equal k_finish's exactly.
//
// Explicit dims (RK_EXPLICIT): every node has its own split dim, so no
// fixed key serves a whole level.  (RK_CHAINED applies the same scheme to
// round-robin subtrees whose roots lie above level D-1.)  Write K(v) for the order node v's elements have
// before v's own sort (the composite key of v's ancestors).  The kernel
// keeps, for every dim d and block v, the order O_d|v = (c[d], K(v)).  The
// subtree root's K is its ancestors' chain, then the input index.  Node v
// with dim dv takes its pivot from O_dv|v; every order is then stably
// partitioned by side, which leaves O_d|child = (c[d], K(v)) while the
// child needs (c[d], c[dv], K(v)) = (c[d], K(child)).  The two differ only
// inside runs of equal c[d], which a stable sort by c[dv] repairs.  Such
// runs come from coordinate ties at the root, marked per slot and dim;
// roots with a tie run longer than RK_RUNMAX are left to the bitonic
// finish.  A node's dim is the widest extent of its domain, found by a
// walk from the subtree root's box down the split table.
// RK_CSTRIDE: one wave's 256 digit counters plus one word, so that the
// lanes of a wave (random digits) and of the scan (same digit, successive
// waves) spread over the LDS banks
// RK_SLOT_BITS: the chained modes keep a node's split dim in its pivot
// word, above the slot (CAP <= 8192).
enum { RK_DMAX = 4, RK_CSTRIDE = 257, RK_RUNMAX = 16, RK_SLOT_BITS = 13 };

template <int CAP, int NT, int DM>
struct RankLDS {
  uint16_t buf[DM + 1][CAP];       // the D <= DM orders plus one scratch, by a buffer map
  uint32_t kbuf[CAP];              // radix key of each slot
  uint32_t cnt[RK_CSTRIDE * (NT / 64)];  // radix: counts at [wave*RK_CSTRIDE + digit]; partition: subtree prefixes
  uint32_t wsum[NT / 64];
  uint32_t rlo[NT / 64], rhi[NT / 64];  // per-wave key bounds of rk_load_keys
  uint16_t pivot[CAP];             // slot (| split dim << RK_SLOT_BITS) of each subtree node, by local BFS index
  uint8_t side[CAP];               // per slot: 0 left, 1 right, 2 pivot
};

// exclusive scan of one value per thread; returns the thread's prefix
template <int NT>
__device__ inline uint32_t rk_block_scan(uint32_t v, uint32_t *wsum)
{
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = (uint32_t)__shfl_up((int)incl, o);
    if (lane >= o) incl += u;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (int i = 0; i < NT / 64; i++) before += i < w ? wsum[i] : 0u;
  __syncthreads();
  return before + incl - v;
}

// One stable LSD pass: dst = src reordered by digit (kbuf[slot] >> sh) & 255.
// Positions >= m hold no element.  Wave w owns positions
// [w*64*E, (w+1)*64*E), visited in (e, lane) order.
template <int CAP, int NT, int DM>
__device__ inline void rk_radix_pass(RankLDS<CAP, NT, DM> &S, const uint16_t *src, uint16_t *dst,
                                     int m, int sh)
{
  enum { E = CAP / NT, NW = NT / 64 };
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = lane; i < 256; i += 64) S.cnt[w * RK_CSTRIDE + i] = 0u;
  // slots and digits first: these loads do not wait on the counters
  uint16_t slot[E];
  uint32_t dig[E], rank[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int pos = w * 64 * E + e * 64 + lane;
    slot[e] = pos < m ? src[pos] : (uint16_t)0;
  }
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int pos = w * 64 * E + e * 64 + lane;
    dig[e] = pos < m ? (S.kbuf[slot[e]] >> sh) & 255u : 256u;
  }
  __syncthreads();
  const uint64_t below = (1ull << lane) - 1ull;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const bool valid = dig[e] < 256u;
    const uint32_t d = dig[e] & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      const uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    const uint32_t before = valid ? S.cnt[w * RK_CSTRIDE + d] : 0u;
    rank[e] = before + (uint32_t)__popcll(peers & below);
    // every peer has read the count before its lowest lane bumps it (the
    // wave's LDS accesses run in order; this keeps the compiler's order)
    __builtin_amdgcn_wave_barrier();
    if (valid && (peers & below) == 0ull) S.cnt[w * RK_CSTRIDE + d] = before + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // exclusive scan of the counts in (digit, wave) order: 4 per thread
  uint32_t c4[4], sum = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int k = 4 * threadIdx.x + i;  // = digit * NW + wave
    c4[i] = S.cnt[(k % NW) * RK_CSTRIDE + k / NW];
    sum += c4[i];
  }
  uint32_t run = rk_block_scan<NT>(sum, S.wsum);
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int k = 4 * threadIdx.x + i;
    S.cnt[(k % NW) * RK_CSTRIDE + k / NW] = run;
    run += c4[i];
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < E; e++)
    if (dig[e] < 256u) dst[S.cnt[w * RK_CSTRIDE + dig[e]] + rank[e]] = slot[e];
  __syncthreads();
}

// kbuf[slot] = the slot's input index (d < 0) or ordered coordinate d.
// Returns how many low key bits vary over the m keys: an LSD sort needs
// no pass over the digits above them (every key shares those).
template <int CAP, int NT, int DM>
__device__ inline int rk_load_keys(RankLDS<CAP, NT, DM> &S, const uint8_t *sdata, const uint32_t *sidx,
                                   int m, int d, const cukd_layout &L, bool isf)
{
  enum { E = CAP / NT, NW = NT / 64 };
  uint32_t v[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = e * NT + threadIdx.x;
    v[e] = i < m ? (d < 0 ? sidx[i] : load_coord_bits(sdata + (size_t)i * L.stride, L, d)) : 0u;
  }
  // the previous keys and bounds may still be read (a sort of no passes
  // has no barrier of its own)
  __syncthreads();
  uint32_t lo = 0xffffffffu, hi = 0u;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = e * NT + threadIdx.x;
    if (i < m) {
      const uint32_t k = d < 0 ? v[e] : ordered_key(v[e], isf);
      S.kbuf[i] = k;
      lo = min(lo, k);
      hi = max(hi, k);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, o));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, o));
  }
  if ((threadIdx.x & 63) == 0) {
    S.rlo[threadIdx.x >> 6] = lo;
    S.rhi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  for (int w = 0; w < NW; w++) {
    lo = min(lo, S.rlo[w]);
    hi = max(hi, S.rhi[w]);
  }
  return lo >= hi ? 0 : 32 - __clz(lo ^ hi);
}

// the buffer holding order d (4 bits per order in 'map')
__device__ inline int rk_buf(uint32_t map, int d) { return (int)((map >> (4 * d)) & 15u); }

// Stable LSD sort of buffer 'a' by the kbuf key, using buffer 't' as
// scratch, skipping digit passes at or above 'bits' (digits every key
// shares leave a stable order unchanged).  Leaves the result in a and the
// other buffer as scratch in t.
template <int CAP, int NT, int DM>
__device__ inline void rk_sort(RankLDS<CAP, NT, DM> &S, int &a, int &t, int m, int bits)
{
  for (int sh = 0; sh < bits; sh += 8) {
    rk_radix_pass(S, S.buf[a], S.buf[t], m, sh);
    const int x = a; a = t; t = x;
  }
}

// Split dim of subtree node 'local' (local BFS index): the widest extent of
// its domain, the root box clamped by the splits on the path from the root
// (as k_finish's walk; widestDimension, box.h:110-125: ties pick the later
// dim).  box0: the root's lower corner, then its upper corner; pv: the
// nodes' pivot words (split dim above the slot bits).
template <int DM>
__device__ inline int rk_node_dim(int local, const uint32_t *box0, const uint32_t *split,
                                  const uint16_t *pv, int D, bool isf)
{
  uint32_t lo[DM], hi[DM];
#pragma unroll
  for (int dd = 0; dd < DM; dd++) {
    lo[dd] = dd < D ? box0[dd] : 0u;
    hi[dd] = dd < D ? box0[D + dd] : 0u;
  }
  while (local > 0) {
    const int parent = (local - 1) >> 1;
    const int pd = pv[parent] >> RK_SLOT_BITS;
    const uint32_t sv = split[parent];
    const bool left = (local & 1) != 0;
#pragma unroll
    for (int dd = 0; dd < DM; dd++) {
      if (dd != pd) continue;
      if (left) {
        const uint32_t u = hi[dd];
        const bool lt = isf ? (__uint_as_float(u) < __uint_as_float(sv)) : ((int)u < (int)sv);
        hi[dd] = lt ? u : sv;
      } else {
        const uint32_t l = lo[dd];
        const bool lt = isf ? (__uint_as_float(sv) < __uint_as_float(l)) : ((int)sv < (int)l);
        lo[dd] = lt ? l : sv;
      }
    }
    local = parent;
  }
  int best = 0;
  if (isf) {
    float wb = 0.f;
#pragma unroll
    for (int dd = 0; dd < DM; dd++) {
      if (dd >= D) break;
      const float w = __uint_as_float(hi[dd]) - __uint_as_float(lo[dd]);
      if (w < wb) continue;
      wb = w;
      best = dd;
    }
  } else {
    long long wb = 0;
#pragma unroll
    for (int dd = 0; dd < DM; dd++) {
      if (dd >= D) break;
      const long long w = (long long)(int)hi[dd] - (long long)(int)lo[dd];
      if (w < wb) continue;
      wb = w;
      best = dd;
    }
  }
  return best;
}

// Chained modes: position q of order T holds an element that ties in
// coordinate d with another element of the subtree (tie bit 2 + d of its
// side byte).  If q starts a run of equal c[d] inside [lo, end), the run is
// stably sorted by coordinate dv (insertion sort: runs are at most
// RK_RUNMAX long).  Other positions of the run return at once.  Runs of one
// order are disjoint, and no run's elements change their c[d], so
// concurrent calls on one order agree on where runs start.
__device__ inline void rk_fix_run(uint16_t *T, const uint8_t *side, int q, int lo, int end, int d,
                                  int dv, const uint8_t *sdata, const cukd_layout &L, bool isf)
{
  auto key = [&](uint16_t sl, int dd) {
    return ordered_key(load_coord_bits(sdata + (size_t)sl * L.stride, L, dd), isf);
  };
  const uint32_t tb = 1u << (2 + d);
  const uint32_t kx = key(T[q], d);
  if (q > lo) {
    const uint16_t y = T[q - 1];
    if ((side[y] & tb) && key(y, d) == kx) return;  // not the run's start
  }
  int len = 1;
  while (q + len < end) {
    const uint16_t y = T[q + len];
    if (!(side[y] & tb) || key(y, d) != kx) break;
    len++;
  }
  for (int i = 1; i < len; i++) {
    const uint16_t x = T[q + i];
    const uint32_t kv = key(x, dv);
    int k = i;
    while (k > 0) {
      const uint16_t y = T[q + k - 1];
      if (!(kv < key(y, dv))) break;
      T[q + k] = y;
      k--;
    }
    T[q + k] = x;
  }
}

// MODE: RK_FIXED (round-robin, finish roots at L0 >= D-1), RK_EXPLICIT
// (explicit dims), RK_CHAINED (round-robin with the explicit-dim machinery:
// any L0, for forests whose small trees are finished from their root).
// st: the subtree states of level L0 (explicit dims); redo: modes other
// than RK_FIXED, one word per workgroup, set to 1 where the bitonic finish
// must build the subtree.
enum { RK_FIXED = 0, RK_EXPLICIT = 1, RK_CHAINED = 2 };
template <int CAP, int NT, int DM, int MODE>
__global__ __launch_bounds__(NT) void k_finish_rank(
    TreeSet<FinTree> set, const uint8_t *dataA, const uint8_t *dataB, const uint32_t *idxA,
    const uint32_t *idxB, uint8_t *out, cukd_layout L, const SegState *st, uint32_t *redo)
{
  enum { E = CAP / NT };
  static_assert(CAP <= (1 << RK_SLOT_BITS), "slots fit below the dim bits");
  // GEN: per-block dims, every order partitioned, tie runs repaired
  constexpr bool GEN = MODE != RK_FIXED, WALK = MODE == RK_EXPLICIT;
  __shared__ RankLDS<CAP, NT, DM> S;
  __shared__ FinTree s_tree;
  __shared__ uint32_t s_box[WALK ? 2 * RK_DMAX : 1];  // domain of the subtree root
  if (threadIdx.x == 0)
    s_tree = find_tree(set, (int)blockIdx.x, [](const FinTree &x) { return x.blk0; });
  __syncthreads();
  const FinTree tr = s_tree;
  const int n = tr.n, L0 = tr.L0;
  const int D = L.num_dims;
  // A finish class that mixes round-robin roots above and below level D-1
  // is launched twice: the chained launch builds the roots above, the
  // fixed one the others (which the chained launch marks done).
  if (MODE == RK_FIXED && L0 < D - 1) return;
  if (MODE == RK_CHAINED && L0 >= D - 1) {
    if (threadIdx.x == 0) redo[blockIdx.x] = 0u;
    return;
  }
  const uint8_t *data = ((L0 & 1) ? dataB : dataA) + (size_t)tr.base * L.stride;
  const uint32_t *idx = ((L0 & 1) ? idxB : idxA) + tr.base;
  out += (size_t)tr.base * L.stride;

  const bool isf = L.scalar_type == CUKD_SCALAR_F32;
  const int h = num_levels_for(n);
  const int hs = h - L0;
  const int Cs = 1 << hs;
  const int s = (int)blockIdx.x - tr.blk0;
  LevelGeom g;
  g.init(n, L0);
  const int g_node = (1 << L0) - 1 + s;
  const int m = g.size(s);
  const int base = L0 > 0 ? g.offset(s) : 0;
  const uint8_t *sdata = data + (size_t)base * L.stride;
  const uint32_t *sidx = idx + base;
  // chained modes: the root's ancestor chain (chain[0] is the root's own
  // dim; the rest, most recent first, then the input index, is the order
  // the subtree arrives in), and for explicit dims the root's domain
  uint32_t chain = 0;
  int clen = 0;
  if (WALK) {
    const SegState &gst = st[tr.seg + s];
    if ((int)threadIdx.x < 2 * D) s_box[threadIdx.x] = gst.box[threadIdx.x];
    if (L0 > 0) {
      chain = gst.chain;
      clen = (int)gst.chain_len;
    }
  } else if (GEN) {
    for (int l = L0; l >= 0 && clen < D; l--) chain |= (uint32_t)(l % D) << (4 * clen++);
  }

  // Fast path: order each dim by its coordinate alone (four passes), then
  // reorder every run of equal coordinates by the rest of its key.  Runs
  // longer than RK_RUNMAX send the subtree to the exact LSD chain below
  // (modes other than RK_FIXED: to the bitonic finish), which also mark
  // the run members: tie bit 2 + d of the slot's side byte.
  // Chained modes also hand over subtrees where more than 1/8 of the
  // elements tie in some dim (coarse grids): the run repairs of every
  // level reload coordinates from memory, and at that density they cost
  // more than the bitonic finish.
  __shared__ int s_long;
  __shared__ uint32_t s_tied[RK_DMAX];  // chained modes: tie-run members per dim
  if (threadIdx.x == 0) s_long = 0;
  if (GEN && threadIdx.x < RK_DMAX) s_tied[threadIdx.x] = 0u;
  bool bail = false;
  uint32_t map = 0;
  int t = DM;
  for (int d = 0; d < D; d++) {
    int a = d;
    for (int i = threadIdx.x; i < m; i += NT) {
      S.buf[a][i] = (uint16_t)i;
      if (GEN && d == 0) S.side[i] = 0;
    }
    rk_sort(S, a, t, m, rk_load_keys(S, sdata, sidx, m, d, L, isf));
    uint16_t *A = S.buf[a];
#pragma unroll
    for (int e = 0; e < E; e++) {
      const int q = threadIdx.x * E + e;
      if (q + 1 >= m) continue;
      const uint32_t kq = S.kbuf[A[q]];
      if (q > 0 && S.kbuf[A[q - 1]] == kq) continue;  // not a run start
      int len = 1;
      while (q + len < m && len <= RK_RUNMAX && S.kbuf[A[q + len]] == kq) len++;
      if (len == 1) continue;
      if (len > RK_RUNMAX) {
        s_long = 1;
        continue;
      }
      // insertion sort of the run by the rest of the key: (c[d-1], ...,
      // c[d-D+1], idx), or (chain[1], ..., idx) in the chained modes
      const int nc = GEN ? clen : D;
      for (int i = 1; i < len; i++) {
        const uint16_t x = A[q + i];
        int k = i;
        while (k > 0) {
          const uint16_t y = A[q + k - 1];
          bool less = false, decided = false;
          for (int c = 1; c < nc && !decided; c++) {
            const int dd = GEN ? (int)chain_at(chain, c) : (d - c + D) % D;
            const uint32_t kx = ordered_key(load_coord_bits(sdata + (size_t)x * L.stride, L, dd), isf);
            const uint32_t ky = ordered_key(load_coord_bits(sdata + (size_t)y * L.stride, L, dd), isf);
            if (kx != ky) {
              less = kx < ky;
              decided = true;
            }
          }
          if (!decided) less = sidx[x] < sidx[y];
          if (!less) break;
          A[q + k] = y;
          k--;
        }
        A[q + k] = x;
      }
      if (GEN) {
        for (int i = 0; i < len; i++) S.side[A[q + i]] |= (uint8_t)(4u << d);
        atomicAdd(&s_tied[d], (uint32_t)len);
      }
    }
    __syncthreads();
    // uniform: both were read after the barrier
    bail = s_long != 0 || (GEN && s_tied[d] * 8u > (uint32_t)m);
    if (bail) break;
    map |= (uint32_t)a << (4 * d);
  }
  if (GEN) {
    if (threadIdx.x == 0) redo[blockIdx.x] = bail ? 1u : 0u;
    if (bail) return;
  } else if (s_long) {
    // exact chain.  key_{D-1}: identity, then stably by idx, c[0], ..., c[D-1]
    int a = 0;
    t = DM;
    for (int i = threadIdx.x; i < m; i += NT) S.buf[a][i] = (uint16_t)i;
    rk_sort(S, a, t, m, rk_load_keys(S, sdata, sidx, m, -1, L, isf));
    for (int d = 0; d < D; d++) rk_sort(S, a, t, m, rk_load_keys(S, sdata, sidx, m, d, L, isf));
    // buffers: order D-1 in a, scratch t, the others free for orders 0..D-2
    map = 0;
    {
      int next = 0;
      for (int d = 0; d + 1 < D; d++) {
        while (next == a || next == t) next++;
        map |= (uint32_t)next << (4 * d);
        next++;
      }
      map |= (uint32_t)a << (4 * (D - 1));
    }
    // key_d = key_{d-1} stably sorted by c[d]
    for (int d = 0; d + 1 < D; d++) {
      int bd = rk_buf(map, d);
      const uint16_t *prev = S.buf[rk_buf(map, d == 0 ? D - 1 : d - 1)];
      for (int i = threadIdx.x; i < m; i += NT) S.buf[bd][i] = prev[i];
      rk_sort(S, bd, t, m, rk_load_keys(S, sdata, sidx, m, d, L, isf));
      map = (map & ~(15u << (4 * d))) | ((uint32_t)bd << (4 * d));
    }
  }

  const int p0 = threadIdx.x * E;
  uint32_t *sizes = S.kbuf;  // per level: subtree size | left-child size << 16, per block
  // chained modes: the split value of every node with grandchildren (local
  // index < CAP/2 - 1), above the per-block words (fewer than CAP/2)
  uint32_t *split = S.kbuf + CAP / 2;
  // Levels whose blocks (B positions) are no wider than one wave's 64 * E
  // positions run wave by wave: every block then lies inside one wave, so
  // the level needs no workgroup barrier (a wave's LDS accesses are
  // performed in order; wave_barrier keeps the compiler's order).  Waves
  // then drift apart by whole levels, and block numbers double per level,
  // so the per-block words (sizes, partition bases) of such levels live in
  // a region of the wave's own: wave w's i-th block at slot w * 32 E + i
  // (a wave holds at most 32 E blocks of >= 2 positions).
  static_assert(32 * E <= RK_CSTRIDE && 32 * E * (NT / 64) <= CAP / 2, "per-wave block slots");
  const int lane = (int)threadIdx.x & 63;
  const int wslot0 = ((int)threadIdx.x >> 6) * 32 * E;
  for (int j = 0; j < hs; j++) {
    const int lgB = hs - j;
    const int B = 1 << lgB;
    const bool wl = B <= 64 * E;  // wave-local level
    const int first_node = (g_node + 1) * (1 << j) - 1;
    const int lvl0 = (1 << j) - 1;  // local index of the level's first node
    const int dj = (L0 + j) % D;
    // slot of block b's words: b itself, or its slot in the wave's region
    const int wb0 = (p0 - lane * E) >> lgB;  // the wave's first block
    auto bslot = [&](int b) { return wl ? wslot0 + (b - wb0) : b; };
    auto block_words = [&](int b, int slot) {
      const int node = first_node + b;
      sizes[slot] = (uint32_t)subtree_size(node, n) | ((uint32_t)subtree_size(2 * node + 1, n) << 16);
      if (WALK)
        S.pivot[lvl0 + b] = (uint16_t)((node < n ? rk_node_dim<DM>(lvl0 + b, s_box, split, S.pivot, D, isf) : 0) << RK_SLOT_BITS);
      else if (GEN)
        S.pivot[lvl0 + b] = (uint16_t)(dj << RK_SLOT_BITS);
    };
    if (wl) {
      if (j == 0 || !(B * 2 <= 64 * E)) __syncthreads();  // leaving the workgroup-wide levels
      for (int i = lane; i < (64 * E) >> lgB; i += 64) {
        const int b = wb0 + i;
        if (b < (1 << j)) block_words(b, wslot0 + i);
      }
      __builtin_amdgcn_wave_barrier();
    } else {
      for (int b = threadIdx.x; b < (1 << j); b += NT) block_words(b, b);
      __syncthreads();
    }
    // pivots and sides from the order of each block's dim (round-robin: the
    // level's dim, whose order is partitioned on the way)
    if (GEN) {
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int p = p0 + e;
        if (p >= Cs) continue;
        const int b = p >> lgB, r = p & (B - 1);
        const uint32_t zs = sizes[bslot(b)];
        if (r >= (int)(zs & 0xffffu)) continue;
        const int left = (int)(zs >> 16);
        const int db = S.pivot[lvl0 + b] >> RK_SLOT_BITS;
        const uint16_t sl = S.buf[rk_buf(map, db)][p];
        S.side[sl] = (uint8_t)((S.side[sl] & 0xfcu) | (r == left ? 2u : (r < left ? 0u : 1u)));
        if (r == left) {
          S.pivot[lvl0 + b] = (uint16_t)(sl | (db << RK_SLOT_BITS));
          if (WALK && B > 2) split[lvl0 + b] = load_coord_bits(sdata + (size_t)sl * L.stride, L, db);
        }
      }
      if (wl) __builtin_amdgcn_wave_barrier();
      else __syncthreads();
    } else {
      const uint16_t *A = S.buf[rk_buf(map, dj)];
      uint16_t *T = S.buf[t];
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int p = p0 + e;
        if (p >= Cs) continue;
        const int b = p >> lgB, r = p & (B - 1);
        const uint32_t zs = sizes[bslot(b)];
        const int sz = (int)(zs & 0xffffu);
        if (r >= sz) continue;
        const int left = (int)(zs >> 16);
        const uint16_t sl = A[p];
        if (r == left) {
          S.side[sl] = 2;
          S.pivot[lvl0 + b] = sl;
        } else {
          S.side[sl] = r < left ? 0 : 1;
          T[r < left ? p : (b << lgB) + (B >> 1) + (r - left - 1)] = sl;
        }
      }
      if (wl) __builtin_amdgcn_wave_barrier();
      else __syncthreads();
      const int x = rk_buf(map, dj);
      map = (map & ~(15u << (4 * dj))) | ((uint32_t)t << (4 * dj));
      t = x;
    }
    if (B <= 2) continue;  // the children are leaves or empty: no later level reads the orders
    // stable partition of the orders by side, inside each subtree (every
    // order in the chained modes; RK_FIXED: the others)
    for (int d = 0; d < D; d++) {
      if (!GEN && d == dj) continue;
      const uint16_t *A = S.buf[rk_buf(map, d)];
      uint16_t *T = S.buf[t];
      uint16_t sl[E];
      uint8_t sd[E];
      uint32_t cntL = 0, cntR = 0, ties = 0;
#pragma unroll
      for (int e = 0; e < E; e++) {
        const int p = p0 + e;
        sd[e] = 3;
        sl[e] = 0;
        if (p < Cs) {
          const int b = p >> lgB, r = p & (B - 1);
          if (r < (int)(sizes[bslot(b)] & 0xffffu)) sl[e] = A[p], sd[e] = 0;
        }
      }
#pragma unroll
      for (int e = 0; e < E; e++) {
        if (sd[e] == 0) {
          const uint32_t v = S.side[sl[e]];
          sd[e] = (uint8_t)(v & 3u);
          if (GEN && sd[e] < 2 && ((v >> (2 + d)) & 1u)) ties |= 1u << e;
        }
        cntL += sd[e] == 0;
        cntR += sd[e] == 1;
      }
      // packed (right << 16 | left) prefix over the positions (of the
      // workgroup, or of the wave on a wave-local level: a block's
      // positions are consecutive either way)
      uint32_t pre;
      const uint32_t packed = (cntR << 16) | cntL;
      if (wl) {
        uint32_t incl = packed;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t u = (uint32_t)__shfl_up((int)incl, o);
          if (lane >= o) incl += u;
        }
        pre = incl - packed;
      } else {
        pre = rk_block_scan<NT>(packed, S.wsum);
      }
      uint32_t at[E];
#pragma unroll
      for (int e = 0; e < E; e++) {
        at[e] = pre;
        pre += (sd[e] == 0) ? 1u : ((sd[e] == 1) ? (1u << 16) : 0u);
        const int p = p0 + e;
        if (p < Cs && (p & (B - 1)) == 0) S.cnt[bslot(p >> lgB)] = at[e];
      }
      if (wl) __builtin_amdgcn_wave_barrier();
      else __syncthreads();
      uint16_t npos[GEN ? E : 1];
#pragma unroll
      for (int e = 0; e < E; e++) {
        if (sd[e] > 1) continue;
        const int p = p0 + e;
        const int b = p >> lgB;
        const uint32_t b0 = S.cnt[bslot(b)];
        const int np = sd[e] == 0 ? (b << lgB) + (int)((at[e] & 0xffffu) - (b0 & 0xffffu))
                                  : (b << lgB) + (B >> 1) + (int)((at[e] >> 16) - (b0 >> 16));
        T[np] = sl[e];
        if (GEN) npos[GEN ? e : 0] = (uint16_t)np;
      }
      if (wl) __builtin_amdgcn_wave_barrier();
      else __syncthreads();
      if (GEN && ties) {
        // runs of equal c[d] in the children: stably by the parent's dim.
        // (Child sizes by closed form: the next level may already be
        // rewriting the per-block words.)
#pragma unroll
        for (int e = 0; e < E; e++) {
          if (!((ties >> e) & 1u)) continue;
          const int b = (p0 + e) >> lgB;
          const int dv = S.pivot[lvl0 + b] >> RK_SLOT_BITS;
          if (dv == d) continue;
          const int node = first_node + b;
          const int right = sd[e];
          const int c0 = (b << lgB) + (right ? (B >> 1) : 0);
          const int csz = subtree_size(2 * node + 1 + right, n);
          rk_fix_run(T, S.side, npos[GEN ? e : 0], c0, c0 + csz, d, dv, sdata, L, isf);
        }
      }
      const int x = rk_buf(map, d);
      map = (map & ~(15u << (4 * d))) | ((uint32_t)t << (4 * d));
      t = x;
    }
  }
  // the nodes, level by level: consecutive local indices are consecutive
  // tree nodes of one level
  __syncthreads();
  for (int local = threadIdx.x; local < Cs - 1; local += NT) {
    const int j = 31 - __clz(local + 1);
    const int node = (g_node + 1) * (1 << j) - 1 + (local - ((1 << j) - 1));
    if (node < n)
      copy_elem(out + (size_t)node * L.stride,
                sdata + (size_t)(S.pivot[local] & ((1u << RK_SLOT_BITS) - 1u)) * L.stride, L,
                GEN ? (int)(S.pivot[local] >> RK_SLOT_BITS) : (L0 + j) % D);
  }
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------

namespace {

inline size_t align_up(size_t v) { return (v + 255) & ~(size_t)255; }

inline int cap_class(int cap) { return 31 - __builtin_clz((unsigned)cap) - 9; }  // 512 -> 0

// Finish-capacity thresholds: a lone tree shrinks its capacity until it
// has this many finish workgroups, and a forest with at least
// CUKD_FOREST_MANY_BLOCKS at the largest capacity keeps it for every tree.
// (The CPU emulator test builds with 1 and 1, so that small inputs reach
// the large capacities.)
#ifndef CUKD_FINISH_MIN_BLOCKS
#define CUKD_FINISH_MIN_BLOCKS 256
#endif
#ifndef CUKD_FOREST_MANY_BLOCKS
#define CUKD_FOREST_MANY_BLOCKS 512
#endif

// LDS-finish capacity and number of top levels of one tree.  In a forest
// with enough finish workgroups to fill the GPU every tree takes the
// largest capacity (fewest top levels); otherwise the capacity shrinks
// until the tree alone has >= 256 finish workgroups.
// Trees that only the bitonic finish can build (explicit dims with D >
// RK_DMAX) start from 4096-element subtrees: its costs grow with log^2 of
// the subtree size, and a 10M float3 explicit-dim build measured 9.6 ms
// against 11.1 ms with 8192 (one more global level).  The rank finish is
// fastest from 8192 (5.0 vs 5.2 ms for round-robin trees).
void tree_cap(int n, bool many, bool bitonic_only, int &cap, int &L0)
{
  const int h = num_levels_for(n);
  L0 = 0;
  for (cap = bitonic_only ? 4096 : 8192; cap >= 512; cap >>= 1) {
    const int lg = 31 - __builtin_clz((unsigned)cap);
    L0 = h - lg > 0 ? h - lg : 0;
    if (many || (1 << L0) >= CUKD_FINISH_MIN_BLOCKS || cap == 512) break;
  }
  if (L0 == 0) {
    // whole tree in one workgroup: capacity = 2^h
    cap = 512;
    while (cap < (1 << h)) cap <<= 1;
  }
}

template <int CAP, int NT>
hipError_t launch_finish(bool expl, const TreeSet<FinTree> &set, int blocks, hipStream_t s,
                         const uint8_t *dA, const uint8_t *dB, const uint32_t *iA,
                         const uint32_t *iB, const SegState *st, uint8_t *out,
                         const cukd_layout &L, const uint32_t *only)
{
  if (expl)
    hipLaunchKernelGGL((k_finish<CAP, NT, true>), dim3(blocks), dim3(NT), 0, s, set, dA, dB, iA,
                       iB, st, out, L, only);
  else
    hipLaunchKernelGGL((k_finish<CAP, NT, false>), dim3(blocks), dim3(NT), 0, s, set, dA, dB, iA,
                       iB, st, out, L, only);
  return hipGetLastError();
}

// Rank finish of one capacity class.  Round-robin: the chained launch for
// finish roots above level D-1 (if any), the fixed one for the others (if
// any).  The chained forms are followed by the bitonic finish of the
// subtrees they flagged in 'redo' (long or dense coordinate tie runs).
template <int CAP, int NT>
hipError_t launch_finish_rank(bool expl, bool any_chained, bool any_fixed,
                              const TreeSet<FinTree> &set, int blocks, hipStream_t s,
                              const uint8_t *dA, const uint8_t *dB, const uint32_t *iA,
                              const uint32_t *iB, const SegState *st, uint8_t *out,
                              const cukd_layout &L, uint32_t *redo)
{
  const bool d3 = L.num_dims <= 3;
#define CUKD_RANK_LAUNCH(MODEv)                                                              \
  do {                                                                                       \
    if (d3)                                                                                  \
      hipLaunchKernelGGL((k_finish_rank<CAP, NT, 3, MODEv>), dim3(blocks), dim3(NT), 0, s,   \
                         set, dA, dB, iA, iB, out, L, st, redo);                             \
    else                                                                                     \
      hipLaunchKernelGGL((k_finish_rank<CAP, NT, RK_DMAX, MODEv>), dim3(blocks), dim3(NT), 0, \
                         s, set, dA, dB, iA, iB, out, L, st, redo);                          \
  } while (0)
  const bool chained = expl || any_chained;
  if (expl) CUKD_RANK_LAUNCH(RK_EXPLICIT);
  else if (any_chained) CUKD_RANK_LAUNCH(RK_CHAINED);
  if (!expl && any_fixed) CUKD_RANK_LAUNCH(RK_FIXED);
#undef CUKD_RANK_LAUNCH
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess || !chained) return e;
  return launch_finish<CAP, NT>(expl, set, blocks, s, dA, dB, iA, iB, st, out, L, redo);
}

}  // namespace

int make_forest_plan(ForestPlan &P, const long long *off, int ntrees, const cukd_layout &L,
                     bool device_tables)
{
  P = ForestPlan();
  P.L = L;
  P.expl = L.dim_offset >= 0;
  P.ntrees = ntrees;
  P.ntotal = off[ntrees] - off[0];
  for (int t = 0; t < ntrees; t++)
    if (off[t + 1] - off[t] > P.max_n) P.max_n = off[t + 1] - off[t];
  P.tables = device_tables;
  const int D = L.num_dims;
  std::vector<int> cap(ntrees, 0), L0(ntrees, 0);
  long long blocks_max_cap = 0;
  for (int t = 0; t < ntrees; t++) {
    const int n = (int)(off[t + 1] - off[t]);
    if (n < 1) continue;
    tree_cap(n, true, P.expl && D > RK_DMAX, cap[t], L0[t]);
    blocks_max_cap += 1ll << L0[t];
  }
  const bool many = blocks_max_cap >= CUKD_FOREST_MANY_BLOCKS;
  P.max_L0 = 0;
  for (int t = 0; t < ntrees; t++) {
    const int n = (int)(off[t + 1] - off[t]);
    if (n < 1) continue;
    tree_cap(n, many, P.expl && D > RK_DMAX, cap[t], L0[t]);
    P.max_L0 = L0[t] > P.max_L0 ? L0[t] : P.max_L0;
  }
  // segment states, numbered level by level
  std::vector<std::vector<int>> seg(ntrees);
  size_t next = 0;
  for (int l = 0; l <= P.max_L0; l++)
    for (int t = 0; t < ntrees; t++) {
      if (off[t + 1] - off[t] < 1 || L0[t] < l) continue;
      seg[t].push_back((int)next);
      next += (size_t)1 << l;
    }
  P.nstates = next;
  // top levels
  P.max_hseg = 1;
  for (int l = 0; l < P.max_L0; l++) {
    std::vector<LvlTree> v;
    int hseg = 0, stl = 0, ptl = 0;
    for (int t = 0; t < ntrees; t++) {
      const int n = (int)(off[t + 1] - off[t]);
      if (n < 1 || L0[t] <= l) continue;
      LvlTree x;
      x.n = n;
      x.base = (int)(off[t] - off[0]);
      x.seg = seg[t][l];
      x.cseg = seg[t][l + 1];
      x.hseg = hseg;
      LevelGeom g;
      g.init(n, l);
      // select tiles: larger than the partition's (fewer, fatter histogram
      // flushes), still at most two subtrees per tile (T <= F/2 <= min size)
      int tsel = 16384;
      while (tsel > g.F / 2) tsel >>= 1;
      x.tsel = tsel;
      x.tpart = cap[t] / 2 < SEL_NT * SEL_ITEMS ? cap[t] / 2 : SEL_NT * SEL_ITEMS;
      x.stile0 = stl;
      stl += (g.n_elems + tsel - 1) / tsel;
      x.ptile0 = ptl;
      ptl += (g.n_elems + x.tpart - 1) / x.tpart;
      x.more = l + 1 < L0[t] ? 1 : 0;
      x.nhseg = 0;
      hseg += 1 << l;
      v.push_back(x);
    }
    P.lvl.push_back(v);
    P.lvl_segs.push_back(hseg);
    P.lvl_stiles.push_back(stl);
    P.lvl_ptiles.push_back(ptl);
    P.max_hseg = (size_t)hseg > P.max_hseg ? (size_t)hseg : P.max_hseg;
  }
  // each tree's slots on the next level (the trees of level l + 1 are
  // those of level l with 'more', in the same order)
  for (int l = 0; l + 1 < P.max_L0; l++) {
    size_t i = 0;
    for (LvlTree &x : P.lvl[l]) {
      if (!x.more) continue;
      while (P.lvl[l + 1][i].base != x.base) i++;
      x.nhseg = P.lvl[l + 1][i].hseg;
    }
  }
  // Histogram copies per subtree where many tiles add into one subtree (at
  // least 64 tiles per copy: the later passes read every copy, and 32 or
  // 16 tiles per copy measured slower at 10M points), as many as the
  // buffers (sized for the level with the most subtrees) hold.
  for (int l = 0; l < P.max_L0; l++) {
    const int segs = P.lvl_segs[l];
    const int tiles_per_seg = P.lvl_stiles[l] / (segs > 0 ? segs : 1);
    int hc = 1;
    while (hc < SEL_HCOPIES && tiles_per_seg >= 64 * hc && (size_t)segs * hc * 2 <= P.max_hseg)
      hc *= 2;
    P.lvl_hcopies.push_back(hc);
  }
  // finish workgroups, by capacity class
  for (int c = 0; c < 5; c++) P.fin_blocks[c] = 0;
  for (int t = 0; t < ntrees; t++) {
    const int n = (int)(off[t + 1] - off[t]);
    if (n < 1) continue;
    const int c = cap_class(cap[t]);
    FinTree f;
    f.n = n;
    f.base = (int)(off[t] - off[0]);
    f.L0 = L0[t];
    f.seg = seg[t][L0[t]];
    f.blk0 = P.fin_blocks[c];
    P.fin_blocks[c] += 1 << L0[t];
    P.fin[c].push_back(f);
  }
  // bounds / copy-in tiles
  P.in_tiles = 0;
  for (int t = 0; t < ntrees; t++) {
    const int n = (int)(off[t + 1] - off[t]);
    if (n < 1) continue;
    InTree x;
    x.n = n;
    x.base = (int)(off[t] - off[0]);
    x.id = t;
    x.seg = seg[t][0];
    x.tile0 = P.in_tiles;
    P.in_tiles += (n + IN_TILE - 1) / IN_TILE;
    P.in.push_back(x);
  }
  // workspace
  const size_t N = (size_t)P.ntotal, S = (size_t)L.stride;
  size_t o = 0;
  P.off_dataA = o; o = align_up(o + N * S);
  P.off_idxA = o; o = align_up(o + N * 4);
  P.off_keyA = o; o = align_up(o + N * 4);
  if (P.max_L0 > 0) {
    P.off_dataB = o; o = align_up(o + N * S);
    P.off_idxB = o; o = align_up(o + N * 4);
    P.off_keyB = o; o = align_up(o + N * 4);
  } else {
    P.off_dataB = P.off_idxB = P.off_keyB = 0;
  }
  P.off_st = o; o = align_up(o + P.nstates * sizeof(SegState));
  P.off_hist = o; o = align_up(o + 3 * P.max_hseg * HIST_STRIDE * 4);  // digits 0, 1, 2
  P.off_stfin = o; o = align_up(o + P.max_hseg * sizeof(SegState));
  P.off_rk = o; o = align_up(o + P.max_hseg * 4 * 4);
  P.off_cursors = o; o = align_up(o + 2 * P.max_hseg * 2 * 4);  // even and odd levels
  P.off_tb = o; o = align_up(o + 2 * P.max_hseg * 2 * MAX_WORDS * 4);  // even and odd levels
  P.zero_end = o;
  P.off_acc = o; o = align_up(o + (size_t)ntrees * 2 * D * 4);
  // explicit-dim rank finish: one flag per finish workgroup (each class
  // launches in turn, so the largest class sets the size)
  int max_fin = 0;
  for (int c = 0; c < 5; c++) max_fin = P.fin_blocks[c] > max_fin ? P.fin_blocks[c] : max_fin;
  P.off_redo = o; o = align_up(o + (size_t)max_fin * 4);
  if (device_tables) {
    P.off_tab_in = o; o = align_up(o + P.in.size() * sizeof(InTree));
    for (int c = 0; c < 5; c++) { P.off_tab_fin[c] = o; o = align_up(o + P.fin[c].size() * sizeof(FinTree)); }
    P.off_tab_lvl.resize(P.lvl.size());
    for (size_t l = 0; l < P.lvl.size(); l++) {
      P.off_tab_lvl[l] = o;
      o = align_up(o + P.lvl[l].size() * sizeof(LvlTree));
    }
  }
  P.total = o;
  return CUKD_OK;
}

// host image of the device tables of a forest plan (offsets relative to
// the workspace)
void forest_tables_image(const ForestPlan &P, std::vector<uint8_t> &img, size_t &begin)
{
  begin = P.off_tab_in;
  img.assign(P.total - begin, 0);
  auto put = [&](size_t off, const void *src, size_t bytes) {
    if (bytes) memcpy(img.data() + (off - begin), src, bytes);
  };
  put(P.off_tab_in, P.in.data(), P.in.size() * sizeof(InTree));
  for (int c = 0; c < 5; c++) put(P.off_tab_fin[c], P.fin[c].data(), P.fin[c].size() * sizeof(FinTree));
  for (size_t l = 0; l < P.lvl.size(); l++)
    put(P.off_tab_lvl[l], P.lvl[l].data(), P.lvl[l].size() * sizeof(LvlTree));
}

template <class E>
static TreeSet<E> make_set(const ForestPlan &P, uint8_t *ws, size_t off, const std::vector<E> &v)
{
  TreeSet<E> s;
  s.count = (int)v.size();
  if (P.tables) {
    s.tab = (const E *)(ws + off);
    memset(&s.one, 0, sizeof(E));
  } else {
    s.tab = nullptr;
    s.one = v[0];
  }
  return s;
}

// Launches a whole build of plan P over d_points with workspace ws (tables
// already in place when P.tables).  No allocation, no host synchronisation.
int forest_build_run(const ForestPlan &P, void *d_points, void *d_bounds, uint8_t *ws,
                     hipStream_t s)
{
  const cukd_layout &L = P.L;
  const bool expl = P.expl;
  const int D = L.num_dims;
  if (P.in.empty()) return CUKD_OK;
  uint8_t *dataA = ws + P.off_dataA, *dataB = ws + P.off_dataB;
  uint32_t *idxA = (uint32_t *)(ws + P.off_idxA), *idxB = (uint32_t *)(ws + P.off_idxB);
  uint32_t *keyA = (uint32_t *)(ws + P.off_keyA), *keyB = (uint32_t *)(ws + P.off_keyB);
  SegState *st = (SegState *)(ws + P.off_st);
  // digit histograms 0, 1, 2; the cursors and tie-word bounds alternate
  // between even and odd levels (a level's tie launch starts the next
  // level's while its own are still in use)
  uint32_t *h0 = (uint32_t *)(ws + P.off_hist);
  uint32_t *h1 = h0 + P.max_hseg * HIST_STRIDE;
  uint32_t *h2 = h1 + P.max_hseg * HIST_STRIDE;
  SegState *stfin = (SegState *)(ws + P.off_stfin);
  uint32_t *rk = (uint32_t *)(ws + P.off_rk);
  uint32_t *cursors = (uint32_t *)(ws + P.off_cursors);
  uint32_t *tbp[2] = {(uint32_t *)(ws + P.off_tb), nullptr};
  tbp[1] = tbp[0] + P.max_hseg * 2 * MAX_WORDS;
  uint32_t *curp[2] = {cursors, cursors + P.max_hseg * 2};
  uint32_t *acc = (uint32_t *)(ws + P.off_acc);
  uint32_t *redo = (uint32_t *)(ws + P.off_redo);
  uint8_t *pts = (uint8_t *)d_points;

  const TreeSet<InTree> in = make_set(P, ws, P.off_tab_in, P.in);
  if (d_bounds || expl) {
    const int nacc = P.ntrees * 2 * D;
    hipLaunchKernelGGL(k_bounds_init, dim3((nacc + 255) / 256), dim3(256), 0, s, acc, D,
                       P.ntrees);
    hipLaunchKernelGGL(k_bounds, dim3(P.in_tiles), dim3(256), 0, s, in, acc,
                       (const uint8_t *)pts, L);
    hipLaunchKernelGGL(k_bounds_final, dim3(((int)P.in.size() + 255) / 256), dim3(256), 0, s, in,
                       (int)P.in.size(), (const uint32_t *)acc, (uint32_t *)d_bounds, st, L,
                       expl ? 1 : 0);
  }
  hipLaunchKernelGGL(k_copy_in, dim3(P.in_tiles), dim3(256), 0, s, in, dataA, idxA,
                     P.max_L0 > 0 ? keyA : (uint32_t *)nullptr, (const uint8_t *)pts, L,
                     (const SegState *)st, expl ? 1 : 0, h0,
                     P.max_L0 > 0 ? (P.zero_end - P.off_hist) / 4 : 0);
  if (P.max_L0 > 0) {
    const TreeSet<LvlTree> set0 = make_set(P, ws, P.tables ? P.off_tab_lvl[0] : 0, P.lvl[0]);
    hipLaunchKernelGGL(k_level_init, dim3((P.lvl_segs[0] + 255) / 256), dim3(256), 0, s, set0,
                       P.lvl_segs[0], st, curp[0], tbp[0], 0, D, expl ? 1 : 0);
  }
  uint8_t *dcur = dataA, *dnext = dataB;
  uint32_t *icur = idxA, *inext = idxB, *kcur = keyA, *knext = keyB;
  for (int level = 0; level < P.max_L0; level++) {
    const TreeSet<LvlTree> set =
        make_set(P, ws, P.tables ? P.off_tab_lvl[level] : 0, P.lvl[level]);
    const int par = level & 1;
    const int hcopies = P.lvl_hcopies[level];
    // key word 0 (the split coordinate) over the whole level, one digit
    // per launch; then the last pick and the few tied subtrees (over their
    // tie groups alone) in one launch, which also starts the next level's
    // subtrees
    for (int d = 0; d < 3; d++)
      hipLaunchKernelGGL(k_select_digit, dim3(P.lvl_stiles[level]), dim3(SELECT_NT), 0, s, set,
                         level, d, (const uint32_t *)kcur, (const SegState *)st, h0, h1, h2, rk,
                         hcopies);
    hipLaunchKernelGGL(k_tie_resolve, dim3(P.lvl_stiles[level]), dim3(SELECT_NT), 0, s, set,
                       level, (const uint32_t *)kcur, (const uint8_t *)dcur,
                       (const uint32_t *)icur, st, stfin, (const uint32_t *)h2,
                       (const uint32_t *)rk, hcopies, dnext, tbp[par], curp[par ^ 1],
                       tbp[par ^ 1], L, expl ? 1 : 0);
    const int W = L.stride >> 2;
    auto kp = W <= 3 ? k_partition<3> : (W == 4 ? k_partition<4> : k_partition<0>);
    hipLaunchKernelGGL(kp, dim3(P.lvl_ptiles[level]), dim3(SEL_NT), 0, s, set, level,
                       (const uint32_t *)kcur, (const uint8_t *)dcur, (const uint32_t *)icur,
                       (const SegState *)st, (const SegState *)stfin, dnext, inext, knext,
                       curp[par], pts, L, expl ? 1 : 0, h0, h1, h2,
                       P.lvl_segs[level] * hcopies);
    uint8_t *t8 = dcur; dcur = dnext; dnext = t8;
    uint32_t *t32 = icur; icur = inext; inext = t32;
    t32 = kcur; kcur = knext; knext = t32;
  }
  hipError_t e = hipGetLastError();
  for (int c = 0; c < 5 && e == hipSuccess; c++) {
    if (P.fin_blocks[c] == 0) continue;
    const TreeSet<FinTree> fs = make_set(P, ws, P.tables ? P.off_tab_fin[c] : 0, P.fin[c]);
    const int b = P.fin_blocks[c];
    // the rank finish for D <= RK_DMAX: fixed orders for round-robin
    // finish roots that see all D dims above them, the chained form for
    // the others and for explicit dims
    int min_L0 = 1 << 30, max_L0 = -1;
    for (const FinTree &f : P.fin[c]) {
      min_L0 = f.L0 < min_L0 ? f.L0 : min_L0;
      max_L0 = f.L0 > max_L0 ? f.L0 : max_L0;
    }
    const bool chained = min_L0 < D - 1, fixed = max_L0 >= D - 1;
    if (D <= RK_DMAX) {
      switch (c) {
        case 0: e = launch_finish_rank<512, 256>(expl, chained, fixed, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, redo); break;
        case 1: e = launch_finish_rank<1024, 256>(expl, chained, fixed, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, redo); break;
        case 2: e = launch_finish_rank<2048, 512>(expl, chained, fixed, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, redo); break;
        case 3: e = launch_finish_rank<4096, 512>(expl, chained, fixed, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, redo); break;
        default: e = launch_finish_rank<8192, 1024>(expl, chained, fixed, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, redo); break;
      }
      continue;
    }
    switch (c) {
      case 0: e = launch_finish<512, 256>(expl, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, nullptr); break;
      case 1: e = launch_finish<1024, 256>(expl, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, nullptr); break;
      case 2: e = launch_finish<2048, 512>(expl, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, nullptr); break;
      case 3: e = launch_finish<4096, 512>(expl, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, nullptr); break;
      default: e = launch_finish<8192, 1024>(expl, fs, b, s, dataA, dataB, idxA, idxB, st, pts, L, nullptr); break;
    }
  }
  if (e != hipSuccess) {
    set_last_error("builder kernel launch failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

// World bounds only (cukd_compute_bounds): the caller's box doubles as the
// accumulator, so no temporary memory is needed.
int build_bounds_only(void *d_bounds, const void *d_points, int n, const cukd_layout &L,
                      hipStream_t s)
{
  uint32_t *acc = (uint32_t *)d_bounds;
  TreeSet<InTree> one;
  one.tab = nullptr;
  one.count = 1;
  one.one.n = n;
  one.one.base = 0;
  one.one.id = 0;
  one.one.seg = 0;
  one.one.tile0 = 0;
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, s, acc, L.num_dims, 1);
  hipLaunchKernelGGL(k_bounds, dim3((n + IN_TILE - 1) / IN_TILE), dim3(256), 0, s, one, acc,
                     (const uint8_t *)d_points, L);
  hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, s, one, 1, (const uint32_t *)acc, acc,
                     (SegState *)nullptr, L, 0);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_last_error("bounds kernel launch failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  return CUKD_OK;
}

size_t build_temp_bytes(int n, const cukd_layout &L)
{
  ForestPlan P;
  const long long off[2] = {0, n};
  make_forest_plan(P, off, 1, L, false);
  return P.total;
}

int build_tree_impl(void *d_points, int n, const cukd_layout &L, void *d_bounds, hipStream_t s,
                    const MemResource &mem)
{
  ForestPlan P;
  const long long off[2] = {0, n};
  make_forest_plan(P, off, 1, L, false);
  uint8_t *ws = nullptr;
  if (mem.alloc((void **)&ws, P.total, s) != 0) return CUKD_ERR_ALLOC;
  const int rc = forest_build_run(P, d_points, d_bounds, ws, s);
  mem.release(ws, s);
  return rc;
}

}  // namespace cukd_amd
