// The CPU kernel emulator's runtime: a pool of host threads that runs one
// workgroup at a time, and the HIP runtime API over host memory (test
// infrastructure; see include/hip/hip_runtime.h).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdlib>

thread_local emu_uint3 threadIdx, blockIdx;
emu_uint3 blockDim, gridDim;

// Fiber switch for x86-64 (System V): saves the callee-saved registers on
// the current stack, stores its stack pointer in *save_sp and continues on
// new_sp.  A plain ucontext swap would also make a signal-mask system call
// on every switch.
extern "C" void emu_switch(void **save_sp, void *new_sp);
asm(R"(
    .text
    .globl emu_switch
    .type emu_switch,@function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
    .size emu_switch, .-emu_switch
)");

namespace emu {

namespace {

struct Fiber {
  void *sp = nullptr;
  int op = OP_NONE;     // sync point it waits at (OP_NONE: runnable)
  bool done = false;
  uint64_t val = 0;     // value deposited at a wave intrinsic
  int arg = 0;          // shuffle source lane
  uint64_t res = 0;     // result of the last wave intrinsic
  unsigned tid = 0;
};

struct Run {
  std::vector<Fiber> f;
  void *sched_sp = nullptr;
  int cur = -1;
  int nt = 0;
  const void *sched_bottom = nullptr;  // sanitizer fiber bookkeeping
  size_t sched_size = 0;
  const std::function<void()> *body = nullptr;
  unsigned bx = 0, by = 0, bz = 0;
  dim3 block;
  uint8_t *dyn = nullptr;
};

Run *g_run = nullptr;
enum { STACK_BYTES = 512 << 10 };

// fiber stacks, reused by every launch, released at exit
struct Stacks {
  std::vector<uint8_t *> s;
  ~Stacks()
  {
    for (uint8_t *p : s) free(p);
  }
} g_stacks;

#if defined(__SANITIZE_ADDRESS__)
extern "C" void __sanitizer_start_switch_fiber(void **fake_stack_save, const void *bottom, size_t size);
extern "C" void __sanitizer_finish_switch_fiber(void *fake_stack_save, const void **bottom_old,
                                                size_t *size_old);
#define EMU_START_SWITCH(save, bottom, size) __sanitizer_start_switch_fiber(save, bottom, size)
#define EMU_FINISH_SWITCH(save, bo, so) __sanitizer_finish_switch_fiber(save, bo, so)
#else
#define EMU_START_SWITCH(save, bottom, size) ((void)0)
#define EMU_FINISH_SWITCH(save, bo, so) ((void)0)
#endif

// back to the scheduler from the current fiber
void yield_to_sched()
{
  Run &r = *g_run;
  void *fake = nullptr;
  EMU_START_SWITCH(&fake, r.sched_bottom, r.sched_size);
  emu_switch(&r.f[r.cur].sp, r.sched_sp);
  EMU_FINISH_SWITCH(fake, nullptr, nullptr);
}

void fiber_main()
{
  Run &r = *g_run;
  EMU_FINISH_SWITCH(nullptr, &r.sched_bottom, &r.sched_size);
  Fiber &me = r.f[r.cur];
  threadIdx.x = me.tid % r.block.x;
  threadIdx.y = (me.tid / r.block.x) % r.block.y;
  threadIdx.z = me.tid / (r.block.x * r.block.y);
  blockIdx.x = r.bx; blockIdx.y = r.by; blockIdx.z = r.bz;
  (*r.body)();
  me.done = true;
  // a finished fiber is never resumed: its fake stack is released
  EMU_START_SWITCH(nullptr, r.sched_bottom, r.sched_size);
  void *dead = nullptr;
  emu_switch(&dead, r.sched_sp);
}

// a fresh fiber's stack: six zero registers, then fiber_main as the
// return address, placed so that fiber_main starts with the stack aligned
// as after a call
void *fresh_stack(uint8_t *base)
{
  uint8_t *top = base + STACK_BYTES;
  uintptr_t a = ((uintptr_t)top - 64) & ~(uintptr_t)15;
  void **slot = (void **)a;
  slot[0] = (void *)&fiber_main;
  void **sp = slot - 6;
  for (int i = 0; i < 6; i++) sp[i] = nullptr;
  return (void *)sp;
}

// runs fiber i until it reaches a sync point or ends
void resume(Run &r, int i)
{
  r.cur = i;
  void *fake = nullptr;
  EMU_START_SWITCH(&fake, g_stacks.s[i], STACK_BYTES);
  emu_switch(&r.sched_sp, r.f[i].sp);
  EMU_FINISH_SWITCH(fake, nullptr, nullptr);
  r.cur = -1;
}

}  // namespace

int lane() { return (int)(g_run->f[g_run->cur].tid & 63); }

uint64_t alive_mask()
{
  Run &r = *g_run;
  const int w0 = (int)(r.f[r.cur].tid & ~63u);
  uint64_t m = 0;
  for (int l = 0; l < 64 && w0 + l < r.nt; l++)
    if (!r.f[w0 + l].done) m |= 1ull << l;
  return m;
}

void sync_point(int op, uint64_t v, int arg)
{
  Run &r = *g_run;
  Fiber &me = r.f[r.cur];
  me.op = op;
  me.val = v;
  me.arg = arg;
  const emu_uint3 keep = threadIdx;
  yield_to_sched();
  threadIdx = keep;
}

uint64_t sync_result() { return g_run->f[g_run->cur].res; }

void launch_impl(dim3 grid, dim3 block, size_t lds, const std::function<void()> &body)
{
  const int nt = (int)(block.x * block.y * block.z);
  gridDim.x = grid.x; gridDim.y = grid.y; gridDim.z = grid.z;
  blockDim.x = block.x; blockDim.y = block.y; blockDim.z = block.z;
  std::vector<uint8_t> dyn(lds + 16);
  while ((int)g_stacks.s.size() < nt) g_stacks.s.push_back((uint8_t *)aligned_alloc(4096, STACK_BYTES));
  Run r;
  r.nt = nt;
  r.body = &body;
  r.block = block;
  r.dyn = dyn.data();
  Run *outer = g_run;
  g_run = &r;
  const int nw = (nt + 63) / 64;
  for (unsigned bz = 0; bz < grid.z; bz++)
    for (unsigned by = 0; by < grid.y; by++)
      for (unsigned bx = 0; bx < grid.x; bx++) {
        r.bx = bx; r.by = by; r.bz = bz;
        r.f.assign(nt, Fiber());
        for (int i = 0; i < nt; i++) {
          r.f[i].tid = (unsigned)i;
          r.f[i].sp = fresh_stack(g_stacks.s[i]);
        }
        for (;;) {
          // run every runnable fiber to its next sync point
          bool ran = false;
          for (int i = 0; i < nt; i++)
            if (!r.f[i].done && r.f[i].op == OP_NONE) {
              resume(r, i);
              ran = true;
            }
          if (ran) continue;
          int live = 0, at_barrier = 0;
          for (int i = 0; i < nt; i++)
            if (!r.f[i].done) {
              live++;
              at_barrier += r.f[i].op == OP_BARRIER;
            }
          if (live == 0) break;
          bool released = false;
          // a wave whose live lanes all wait at one wave intrinsic.  With
          // EMU_SKEW=1 (2) only the lowest (highest) such wave is released,
          // so that it runs on to the next workgroup barrier before any
          // other wave moves: races between waves that drift apart show.
          static const int skew = getenv("EMU_SKEW") ? atoi(getenv("EMU_SKEW")) : 0;
          for (int wi = 0; wi < nw && !(skew && released); wi++) {
            const int w = skew == 2 ? nw - 1 - wi : wi;
            int op = -1;
            bool all = true, any = false;
            uint64_t mask = 0;
            for (int l = 0; l < 64 && 64 * w + l < nt; l++) {
              const Fiber &f = r.f[64 * w + l];
              if (f.done) continue;
              any = true;
              if (f.op == OP_BARRIER || f.op == OP_NONE || (op >= 0 && f.op != op)) { all = false; break; }
              op = f.op;
              if (f.val) mask |= 1ull << l;
            }
            if (!any || !all) continue;
            for (int l = 0; l < 64 && 64 * w + l < nt; l++) {
              Fiber &f = r.f[64 * w + l];
              if (f.done) continue;
              if (op == OP_BALLOT) f.res = mask;
              else if (op == OP_SHFL) f.res = r.f[64 * w + f.arg].val;
              f.op = OP_NONE;
            }
            released = true;
          }
          if (released) continue;
          if (at_barrier == live) {
            for (int i = 0; i < nt; i++)
              if (!r.f[i].done) r.f[i].op = OP_NONE;
            continue;
          }
          fprintf(stderr, "emulator: workgroup %u deadlocked (lanes of a wave wait at "
                          "different sync points)\n", bx);
          abort();
        }
      }
  g_run = outer;
}

}  // namespace emu

void *emu_dynamic_lds() { return emu::g_run->dyn; }

// ---- runtime API ----
const char *hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "emulated error"; }
hipError_t hipGetLastError(void) { return hipSuccess; }
hipError_t hipMalloc(void **p, size_t n)
{
  *p = aligned_alloc(256, (n + 255) / 256 * 256 + 256);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipMallocAsync(void **p, size_t n, hipStream_t) { return hipMalloc(p, n); }
hipError_t hipMallocManaged(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostMalloc(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipFree(void *p) { free(p); return hipSuccess; }
hipError_t hipFreeAsync(void *p, hipStream_t) { free(p); return hipSuccess; }
hipError_t hipHostFree(void *p) { free(p); return hipSuccess; }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { if (n) memmove(d, s, n); return hipSuccess; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t)
{
  return hipMemcpy(d, s, n, k);
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) { memset(d, v, n); return hipSuccess; }
hipError_t hipMemcpyToSymbol(const void *sym, const void *src, size_t n, size_t off, hipMemcpyKind)
{
  memcpy((uint8_t *)sym + off, src, n);
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = (void *)sym; return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (hipStream_t)new int(0); return hipSuccess; }
hipError_t hipStreamDestroy(hipStream_t s) { delete (int *)s; return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus *st) { *st = hipStreamCaptureStatusNone; return hipSuccess; }
hipError_t hipDeviceSynchronize(void) { return hipSuccess; }
struct emu_event { std::chrono::steady_clock::time_point t; };
hipError_t hipEventCreate(hipEvent_t *e) { *e = new emu_event(); return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = std::chrono::steady_clock::now(); return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b)
{
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int *v, hipDeviceAttribute_t, int) { *v = 4; return hipSuccess; }
