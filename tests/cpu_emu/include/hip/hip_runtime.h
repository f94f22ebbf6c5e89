// hip/hip_runtime.h for the CPU kernel emulator (test infrastructure).
//
// Compiles the library's unmodified .hip sources with g++ so that they can
// run under AddressSanitizer / UndefinedBehaviorSanitizer on the host:
//  * a kernel launch runs its workgroups one after another; the threads of
//    a workgroup are fibers on one host thread, so __shared__ variables
//    (function statics here) are shared exactly by one workgroup's threads;
//  * __syncthreads waits for every thread of the workgroup, and every wave
//    intrinsic (ballot, shuffles, any, wave_barrier) is a rendezvous of the
//    wave's 64 lanes; a thread that returns no longer counts, as a
//    finished lane on the GPU;
//  * device memory is host memory, streams run synchronously, and
//    atomics are host atomics (memory scopes ignored).
// Kernels whose wave intrinsics are reached by only some lanes of a wave
// (divergent control flow) cannot be emulated; the library's are all
// reached by whole waves.
#pragma once

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "hip/hip_runtime_api.h"

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __constant__
#define __shared__ static
#define __launch_bounds__(...)
#define __align__(n) alignas(n)
#define __restrict__

struct emu_uint3 {
  unsigned x = 0, y = 0, z = 0;
};
extern thread_local emu_uint3 threadIdx, blockIdx;
extern emu_uint3 blockDim, gridDim;

namespace emu {

// Sync points of the running workgroup.  Every thread of a workgroup is a
// fiber on one host thread; a fiber runs until it reaches a barrier, a
// wave intrinsic or its end, so the emulation is deterministic and free of
// host data races.
enum Op { OP_NONE = 0, OP_BARRIER, OP_BALLOT, OP_SHFL, OP_WAVE_BARRIER };
void sync_point(int op, uint64_t v, int arg);   // waits until released
uint64_t sync_result();                         // this lane's result
int lane();
uint64_t alive_mask();

inline uint64_t ballot(bool p)
{
  sync_point(OP_BALLOT, p ? 1u : 0u, 0);
  return sync_result();
}
inline uint64_t wave_read(uint64_t v, int src)
{
  sync_point(OP_SHFL, v, src & 63);
  return sync_result();
}

void launch_impl(dim3 grid, dim3 block, size_t lds, const std::function<void()> &body);

template <typename... P, typename... A>
inline void launch(void (*k)(P...), dim3 grid, dim3 block, size_t lds, A &&...args)
{
  std::function<void()> body = [&]() { k(args...); };
  launch_impl(grid, block, lds, body);
}

}  // namespace emu

#define hipLaunchKernelGGL(k, grid, block, lds, stream, ...) \
  emu::launch(k, dim3(grid), dim3(block), (size_t)(lds), ##__VA_ARGS__)

inline void __syncthreads() { emu::sync_point(emu::OP_BARRIER, 0, 0); }
inline void __builtin_amdgcn_wave_barrier() { emu::sync_point(emu::OP_WAVE_BARRIER, 0, 0); }
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(order)
#define __builtin_amdgcn_s_barrier() __syncthreads()

inline uint64_t __ballot(int p) { return emu::ballot(p != 0); }
inline int __any(int p) { return emu::ballot(p != 0) != 0; }
inline int __all(int p) { return emu::ballot(p != 0) == emu::alive_mask(); }
template <typename T>
inline T __shfl(T v, int src, int width = 64)
{
  (void)width;
  uint64_t bits = 0;
  memcpy(&bits, &v, sizeof(T));
  const uint64_t r = emu::wave_read(bits, src);
  T out;
  memcpy(&out, &r, sizeof(T));
  return out;
}
template <typename T>
inline T __shfl_up(T v, unsigned d, int width = 64)
{
  (void)width;
  const int l = emu::lane();
  return __shfl(v, l >= (int)d ? l - (int)d : l);
}
template <typename T>
inline T __shfl_xor(T v, int m, int width = 64)
{
  (void)width;
  return __shfl(v, emu::lane() ^ m);
}
inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned mask, unsigned acc)
{
  const int l = emu::lane();
  const unsigned below = l >= 32 ? 0xffffffffu : ((1u << l) - 1u);
  return acc + (unsigned)__builtin_popcount(mask & below);
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned mask, unsigned acc)
{
  const int l = emu::lane();
  const unsigned below = l < 32 ? 0u : (l == 32 ? 0u : ((1u << (l - 32)) - 1u));
  return acc + (unsigned)__builtin_popcount(mask & below);
}

// bit casts and integer intrinsics
inline float __uint_as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline uint32_t __float_as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
inline float __int_as_float(int32_t u) { float f; memcpy(&f, &u, 4); return f; }
inline int32_t __float_as_int(float f) { int32_t u; memcpy(&u, &f, 4); return u; }
inline int __clz(int x) { return x == 0 ? 32 : __builtin_clz((unsigned)x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline unsigned __umul24(unsigned a, unsigned b) { return (a & 0xffffffu) * (b & 0xffffffu); }
inline unsigned __umulhi(unsigned a, unsigned b) { return (unsigned)(((uint64_t)a * b) >> 32); }
using std::max;
using std::min;

// atomics (memory scopes ignored: host atomics are coherent)
#define __HIP_MEMORY_SCOPE_WAVEFRONT 0
#define __HIP_MEMORY_SCOPE_WORKGROUP 1
#define __HIP_MEMORY_SCOPE_AGENT 2
#define __HIP_MEMORY_SCOPE_SYSTEM 3
#define __hip_atomic_load(p, order, scope) __atomic_load_n(p, order)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n(p, v, order)
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add(p, v, order)
template <typename T> inline T atomicAdd(T *p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline float atomicAdd(float *p, float v)
{
  float old = *p, nv;
  do { nv = old + v; } while (!__atomic_compare_exchange(p, &old, &nv, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST));
  return old;
}
template <typename T> inline T atomicOr(T *p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <typename T> inline T atomicAnd(T *p, T v) { return __atomic_fetch_and(p, v, __ATOMIC_SEQ_CST); }
template <typename T> inline T atomicExch(T *p, T v) { return __atomic_exchange_n(p, v, __ATOMIC_SEQ_CST); }
template <typename T> inline T atomicMin(T *p, T v)
{
  T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}
template <typename T> inline T atomicMax(T *p, T v)
{
  T old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return old;
}

// dynamic LDS of the running workgroup
void *emu_dynamic_lds();
#define HIP_DYNAMIC_SHARED(type, var) type *var = (type *)emu_dynamic_lds();
