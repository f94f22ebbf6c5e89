// cukd/knn.h -- k-nearest-neighbour queries (drop-in for the reference's
// cukd/knn.h): the candidate lists, stackBased / stackFree / cct ::knn as
// device-inline functions, and a batched host call (cukd::batch::knn).
//
// A candidate is the 64-bit key (float bits of dist2) << 32 | uint32(id):
// lists keep the k smallest keys, so equal distances resolve to the
// smaller id, and "no candidate" is (cutOffRadius^2, 0xffffffff).
#pragma once

#include "cukd/fcp.h"

namespace cukd {

struct BaseCandidateList {
  static inline __both__ uint64_t encode(float f, int i)
  {
    union { float f; uint32_t u; } c;
    c.f = f;
    return ((uint64_t)c.u << 32) | (uint32_t)i;
  }
  static inline __both__ float decode_dist2(uint64_t v)
  {
    union { float f; uint32_t u; } c;
    c.u = (uint32_t)(v >> 32);
    return c.f;
  }
  static inline __both__ int decode_pointID(uint64_t v) { return (int)(uint32_t)v; }
};

template <int k>
struct CandidateList : public BaseCandidateList {
  inline __both__ CandidateList(float /*cutOffRadius*/) {}
  inline __both__ float get_dist2(int i) const { return decode_dist2(entry[i]); }
  inline __both__ int get_pointID(int i) const { return decode_pointID(entry[i]); }
  uint64_t entry[k];
  enum { num_k = k };
};

// sorted ascending; insertion is k unrolled compare-exchanges in registers
template <int k>
struct FixedCandidateList : public CandidateList<k> {
  using CandidateList<k>::entry;
  inline __both__ FixedCandidateList(float cutOffRadius) : CandidateList<k>(cutOffRadius)
  {
    const uint64_t v = BaseCandidateList::encode(cutOffRadius * cutOffRadius, -1);
#pragma unroll
    for (int i = 0; i < k; i++) entry[i] = v;
  }
  inline __both__ float maxRadius2() const { return BaseCandidateList::decode_dist2(entry[k - 1]); }
  inline __both__ float returnValue() const { return maxRadius2(); }
  inline __both__ float initialCullDist2() const { return maxRadius2(); }
  inline __both__ void push(float dist, int pointID)
  {
    uint64_t v = BaseCandidateList::encode(dist, pointID);
    // A key at or above the current k-th one would bubble through every
    // slot and fall off the end: skipping the network changes nothing
    // (keys are unique, ids being distinct).
    if (v >= entry[k - 1]) return;
    // insert at the first slot holding a larger key, shifting the rest up
    // by one (the same result as k compare-exchanges, one compare a slot)
    bool gt[k];
#pragma unroll
    for (int i = 0; i < k; i++) gt[i] = entry[i] > v;
    // The selects take an opaque copy of v: otherwise the compiler turns
    // "gt[i] ? v : entry[i]" into min(entry[i], v) and compares every slot
    // twice.
    uint64_t w = v;
#ifdef __HIP_DEVICE_COMPILE__
    asm("" : "+v"(w));
#endif
#pragma unroll
    for (int i = k - 1; i > 0; i--) entry[i] = gt[i - 1] ? entry[i - 1] : (gt[i] ? w : entry[i]);
    entry[0] = gt[0] ? w : entry[0];
  }
  inline __both__ float processCandidate(int candPrimID, float candDist2)
  {
    push(candDist2, candPrimID);
    return maxRadius2();
  }
};

namespace detail {
// max-heap push shared by the heap lists: replace the root, sift down
inline __both__ void heap_push(uint64_t *e, int k, uint64_t v)
{
  if (v >= e[0]) return;
  int pos = 0;
  for (;;) {
    uint64_t largest_val = ~0ull;
    const int c0 = 2 * pos + 1;
    int largest = k;
    if (c0 < k) { largest = c0; largest_val = e[c0]; }
    const int c1 = c0 + 1;
    if (c1 < k && e[c1] > largest_val) { largest = c1; largest_val = e[c1]; }
    if (largest == k || largest_val < v) { e[pos] = v; break; }
    e[pos] = largest_val;
    pos = largest;
  }
}
}  // namespace detail

// max-heap order, maximum at entry[0]
template <int k>
struct HeapCandidateList : public CandidateList<k> {
  using CandidateList<k>::entry;
  inline __both__ HeapCandidateList(float cutOffRadius) : CandidateList<k>(cutOffRadius)
  {
    const uint64_t v = BaseCandidateList::encode(cutOffRadius * cutOffRadius, -1);
    for (int i = 0; i < k; i++) entry[i] = v;
  }
  inline __both__ float maxRadius2() const { return BaseCandidateList::decode_dist2(entry[0]); }
  inline __both__ float returnValue() const { return maxRadius2(); }
  inline __both__ float initialCullDist2() const { return maxRadius2(); }
  inline __both__ void push(float dist, int pointID)
  {
    detail::heap_push(entry, k, BaseCandidateList::encode(dist, pointID));
  }
  inline __both__ float processCandidate(int candPrimID, float candDist2)
  {
    push(candDist2, candPrimID);
    return maxRadius2();
  }
};

// heap list with run-time k over caller-provided storage; a negative radius
// keeps the existing contents (to continue a search across several trees)
struct FlexHeapCandidateList : public BaseCandidateList {
  inline __both__ FlexHeapCandidateList(uint64_t *entryMem, int k, float cutOffRadius)
      : entry(entryMem), k(k)
  {
    if (cutOffRadius >= 0.f) {
      const uint64_t v = encode(cutOffRadius * cutOffRadius, -1);
      for (int i = 0; i < k; i++) entry[i] = v;
    }
  }
  inline __both__ float maxRadius2() const { return decode_dist2(entry[0]); }
  inline __both__ float returnValue() const { return maxRadius2(); }
  inline __both__ float initialCullDist2() const { return maxRadius2(); }
  inline __both__ void push(float dist, int pointID) { detail::heap_push(entry, k, encode(dist, pointID)); }
  inline __both__ float processCandidate(int candPrimID, float candDist2)
  {
    push(candDist2, candPrimID);
    return maxRadius2();
  }
  uint64_t *const entry;
  int const k;
};

namespace stackBased {
template <typename CandidateList, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ float knn(CandidateList &result, typename data_traits::point_t queryPoint,
                          const data_t *d_nodes, int N)
{
  traverse_default<CandidateList, data_t, data_traits>(result, queryPoint, d_nodes, N);
  return result.returnValue();
}
template <typename CandidateList, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ float knn(CandidateList &result, typename data_traits::point_t queryPoint,
                          const box_t<typename data_traits::point_t> /*worldBounds*/,
                          const data_t *d_nodes, int N)
{
  return knn<CandidateList, data_t, data_traits>(result, queryPoint, d_nodes, N);
}
}  // namespace stackBased

namespace stackFree {
template <typename CandidateList, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ float knn(CandidateList &result, typename data_traits::point_t queryPoint,
                          const data_t *d_nodes, int N)
{
  traverse_stack_free<CandidateList, data_t, data_traits>(result, queryPoint, d_nodes, N);
  return result.returnValue();
}
template <typename CandidateList, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ float knn(CandidateList &result, typename data_traits::point_t queryPoint,
                          const box_t<typename data_traits::point_t> /*worldBounds*/,
                          const data_t *d_nodes, int N)
{
  return knn<CandidateList, data_t, data_traits>(result, queryPoint, d_nodes, N);
}
}  // namespace stackFree

namespace cct {
template <typename CandidateList, typename data_t, typename data_traits = default_data_traits<data_t>>
inline __both__ float knn(CandidateList &result, typename data_traits::point_t queryPoint,
                          const box_t<typename data_traits::point_t> worldBounds,
                          const data_t *d_nodes, int N)
{
  traverse_cct<CandidateList, data_t, data_traits>(result, queryPoint, worldBounds, d_nodes, N);
  return result.returnValue();
}
}  // namespace cct

namespace batch {
// For query i: d_ids[i*k+j] / d_dist2[i*k+j] are entry j of its list and
// d_maxRadius2[i] (nullable) is knn()'s return value.
template <typename data_t, typename data_traits = default_data_traits<data_t>>
inline void knn(int traversal, int listKind, int k,
                const typename data_traits::point_t *d_queries, int numQueries,
                const data_t *d_nodes, int numNodes,
                const box_t<typename data_traits::point_t> *d_worldBounds, int *d_ids,
                float *d_dist2, float *d_maxRadius2 = nullptr, float cutOffRadius = INFINITY,
                hipStream_t stream = 0, int flags = 0)
{
  const cukd_layout L = layout_of<data_t, data_traits>();
  detail::check(cukd_knn(traversal, listKind, k, d_nodes, numNodes, &L, d_worldBounds,
                         (const float *)d_queries, numQueries, cutOffRadius, d_ids, d_dist2,
                         d_maxRadius2, flags, stream, nullptr),
                "batch::knn");
}
}  // namespace batch

}  // namespace cukd
