// hip/hip_runtime_api.h for the CPU kernel emulator (test infrastructure):
// the HIP runtime API the library calls, over host memory, with streams
// that run synchronously.
#pragma once

#include <stddef.h>
#include <stdint.h>

typedef int hipError_t;
enum {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNotReady = 600,
};
typedef struct emu_stream *hipStream_t;
typedef struct emu_event *hipEvent_t;
typedef enum { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost,
               hipMemcpyDeviceToDevice, hipMemcpyDefault } hipMemcpyKind;
typedef enum { hipStreamCaptureStatusNone = 0, hipStreamCaptureStatusActive = 1 } hipStreamCaptureStatus;
typedef enum { hipDeviceAttributeMultiprocessorCount = 0 } hipDeviceAttribute_t;
typedef enum { hipFuncAttributeMaxDynamicSharedMemorySize = 0 } hipFuncAttribute;
#define hipStreamNonBlocking 1
#define hipEventDisableTiming 2
#define hipHostMallocDefault 0

struct dim3 {
  unsigned x, y, z;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

struct uint2 { unsigned x, y; };
struct int2 { int x, y; };
struct int3 { int x, y, z; };
struct int4 { int x, y, z, w; };
struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct float4 { float x, y, z, w; };
inline uint2 make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }

#define HIP_SYMBOL(x) (&(x))

const char *hipGetErrorString(hipError_t e);
hipError_t hipGetLastError(void);
hipError_t hipMalloc(void **p, size_t n);
hipError_t hipMallocAsync(void **p, size_t n, hipStream_t s);
hipError_t hipMallocManaged(void **p, size_t n, unsigned flags = 0);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags);
hipError_t hipFree(void *p);
hipError_t hipFreeAsync(void *p, hipStream_t s);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st);
hipError_t hipMemcpyToSymbol(const void *sym, const void *src, size_t n, size_t off = 0,
                             hipMemcpyKind k = hipMemcpyHostToDevice);
hipError_t hipGetSymbolAddress(void **p, const void *sym);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *st);
hipError_t hipDeviceSynchronize(void);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t a, int dev);
template <typename F>
inline hipError_t hipFuncSetAttribute(F, hipFuncAttribute, int) { return hipSuccess; }
template <typename F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int *n, F, int, size_t)
{
  *n = 2;
  return hipSuccess;
}
