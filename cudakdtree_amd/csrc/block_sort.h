// block_sort.h -- workgroup-wide bitonic sort of 64-bit keys with a 32-bit
// payload, shared by the LDS finish kernels of both tree builders.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cukd_amd {

// Register-blocked bitonic sort of CAP 64-bit keys (with a 32-bit payload)
// held E = CAP/NT per thread, thread t owning positions t*E .. t*E+E-1;
// sorts independent aligned blocks of size B ascending.  Strides below E
// stay in registers, strides below 64*E use wave shuffles, larger ones go
// through LDS, 32 bits at a time, so the sort needs only CAP words of LDS.
// Keys are unique apart from the ~0 padding.
template <int CAP, int NT>
struct BlockSort {
  enum { E = CAP / NT };

  // keep min (true) or max of (key, other) at position i
  static __device__ inline void cmpx(uint64_t &k, uint32_t &v, uint64_t ok, uint32_t ov, bool keep_min)
  {
    const bool take = keep_min ? (ok < k) : (ok > k);
    k = take ? ok : k;
    v = take ? ov : v;
  }

  template <int J>
  static __device__ inline void reg_stage(uint64_t (&k)[E], uint32_t (&v)[E], int k_size, int B)
  {
    const int base = threadIdx.x * E;
#pragma unroll
    for (int e = 0; e < E; e++) {
      if (e & J) continue;
      const int i = base + e;
      const bool asc = (k_size == B) || ((i & k_size) == 0);
      const uint64_t a = k[e], b = k[e + J];
      const bool sw = asc ? (a > b) : (a < b);
      if (sw) {
        k[e] = b; k[e + J] = a;
        const uint32_t t = v[e]; v[e] = v[e + J]; v[e + J] = t;
      }
    }
  }

  // sv: CAP 32-bit words of LDS for the cross-wave stages
  static __device__ inline void sort(uint64_t (&k)[E], uint32_t (&v)[E], int B, uint32_t *sv)
  {
    const int base = threadIdx.x * E;
    for (int ks = 2; ks <= B; ks <<= 1) {
      for (int j = ks >> 1; j > 0; j >>= 1) {
        if (j < E) {
          if (j >= 8) { if (E > 8) reg_stage<(E > 8 ? 8 : 1)>(k, v, ks, B); }
          else if (j == 4) { if (E > 4) reg_stage<(E > 4 ? 4 : 1)>(k, v, ks, B); }
          else if (j == 2) { if (E > 2) reg_stage<(E > 2 ? 2 : 1)>(k, v, ks, B); }
          else { reg_stage<1>(k, v, ks, B); }
          continue;
        }
        const int lx = j / E;  // partner thread = tid ^ lx, same e
        if (lx < 64) {
#pragma unroll
          for (int e = 0; e < E; e++) {
            const uint64_t ok = __shfl_xor(k[e], lx);
            const uint32_t ov = __shfl_xor(v[e], lx);
            const int i = base + e;
            const bool asc = (ks == B) || ((i & ks) == 0);
            cmpx(k[e], v[e], ok, ov, ((i & j) == 0) == asc);
          }
        } else {
          // cross-wave stage through one 32-bit LDS word per position: the
          // key's high half, its low half, then the payload
          uint64_t ok[E];
          uint32_t ov[E];
#pragma unroll
          for (int e = 0; e < E; e++) sv[base + e] = (uint32_t)(k[e] >> 32);
          __syncthreads();
#pragma unroll
          for (int e = 0; e < E; e++) ok[e] = (uint64_t)sv[(base + e) ^ j] << 32;
          __syncthreads();
#pragma unroll
          for (int e = 0; e < E; e++) sv[base + e] = (uint32_t)k[e];
          __syncthreads();
#pragma unroll
          for (int e = 0; e < E; e++) ok[e] |= sv[(base + e) ^ j];
          __syncthreads();
#pragma unroll
          for (int e = 0; e < E; e++) sv[base + e] = v[e];
          __syncthreads();
#pragma unroll
          for (int e = 0; e < E; e++) ov[e] = sv[(base + e) ^ j];
          __syncthreads();
#pragma unroll
          for (int e = 0; e < E; e++) {
            const int i = base + e;
            const bool asc = (ks == B) || ((i & ks) == 0);
            cmpx(k[e], v[e], ok[e], ov[e], ((i & j) == 0) == asc);
          }
        }
      }
    }
  }
};

}  // namespace cukd_amd
