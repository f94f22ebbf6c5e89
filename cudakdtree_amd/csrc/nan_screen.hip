// nan_screen.hip -- rejects float point sets that contain NaN coordinates
// before a single-tree build touches them.
//
// The reference's comparator is a plain '<' on coordinates
// (builder_thrust.h:401-404), which is no strict weak order once a NaN is
// present, so its result is undefined there.  cukd_build_tree screens the
// coordinates first and returns CUKD_ERR_INVALID, leaving the array as it
// was.  The result word of each call lives in a library-owned device array
// (one 256-byte line per slot), so no builder, the in-place one included,
// allocates for it.  A free slot's word is zero: it starts so, a scan only
// sets it, and a call that found a NaN writes zero back before it frees
// the slot.
#include <condition_variable>
#include <map>
#include <mutex>

#include "cukd_common.h"

namespace cukd_amd {

enum { NAN_SLOTS = 64, NAN_SLOT_WORDS = 64 };

static __device__ __align__(256) uint32_t g_nan_slots[NAN_SLOTS * NAN_SLOT_WORDS];


// One thread per element (grid-stride): any NaN coordinate sets the flag.
__global__ __launch_bounds__(256) void k_nan_scan(const uint8_t *points, long long n,
                                                  cukd_layout L, uint32_t *flag)
{
  bool bad = false;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const uint8_t *e = points + i * L.stride;
    for (int d = 0; d < L.num_dims; d++) {
      const uint32_t b = load_coord_bits(e, L, d);
      bad |= (b & 0x7f800000u) == 0x7f800000u && (b & 0x007fffffu) != 0u;
    }
  }
  if (__any(bad) && (threadIdx.x & 63) == 0) flag[0] = 1u;
}

namespace {

std::mutex g_mu;
std::condition_variable g_cv;
uint64_t g_busy = 0;                  // slots in use, one bit each
std::map<int, uint32_t *> g_base;     // device -> address of g_nan_slots

// device address of this device's copy of the slot array
int slot_base(uint32_t **base)
{
  int dev = 0;
  CUKD_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_base.find(dev);
  if (it == g_base.end()) {
    void *p = nullptr;
    CUKD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(g_nan_slots)));
    it = g_base.emplace(dev, (uint32_t *)p).first;
  }
  *base = it->second;
  return CUKD_OK;
}

int take_slot()
{
  std::unique_lock<std::mutex> lk(g_mu);
  g_cv.wait(lk, [] { return g_busy != ~0ull; });
  const int s = __builtin_ctzll(~g_busy);
  g_busy |= 1ull << s;
  return s;
}

void give_slot(int s)
{
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_busy &= ~(1ull << s);
  }
  g_cv.notify_one();
}

}  // namespace

// CUKD_OK when no coordinate is NaN (and for int layouts), CUKD_ERR_INVALID
// when one is.  Synchronises 'stream'.  Under stream capture the screen is
// skipped (a capture cannot wait for a result), and CUKD_OK is returned.
int screen_nan(const void *d_points, int n, const cukd_layout &L, hipStream_t stream)
{
  if (L.scalar_type != CUKD_SCALAR_F32 || n < 1) return CUKD_OK;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  CUKD_HIP(hipStreamIsCapturing(stream, &cs));
  if (cs != hipStreamCaptureStatusNone) return CUKD_OK;
  uint32_t *base = nullptr;
  if (int rc = slot_base(&base)) return rc;
  const int slot = take_slot();
  uint32_t *flag = base + slot * NAN_SLOT_WORDS;
  uint32_t h = 0;
  const long long blocks = ((long long)n + 255) / 256;
  hipLaunchKernelGGL(k_nan_scan, dim3((unsigned)(blocks < 4096 ? blocks : 4096)), dim3(256), 0,
                     stream, (const uint8_t *)d_points, (long long)n, L, flag);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, 4, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e == hipSuccess && h) {
    const uint32_t zero = 0u;
    e = hipMemcpyAsync(flag, &zero, 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
  }
  give_slot(slot);
  if (e != hipSuccess) {
    set_last_error("NaN screen failed: %s", hipGetErrorString(e));
    return CUKD_ERR_HIP;
  }
  if (h) {
    set_last_error("points contain NaN coordinates (no strict weak order for the "
                   "builder's '<' comparison); the array is left unchanged");
    return CUKD_ERR_INVALID;
  }
  return CUKD_OK;
}

}  // namespace cukd_amd
