// Runtime of the CPU emulator (see hip/hip_runtime.h): fibers for the threads of a workgroup,
// the wavefront and workgroup collectives, and the HIP host API on host memory.
#include <sanitizer/common_interface_defs.h>
#include <sys/mman.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hip/hip_runtime.h"

extern "C" void emu_ctx_switch(void** save_sp, void* load_sp);
extern "C" void emu_fiber_entry();

namespace emu {
namespace {

constexpr size_t kStack = 1u << 20;  // per fiber (AddressSanitizer frames are large)
constexpr int kPayload = 64;         // bytes of one lane's collective payload

enum State { kReady = 0, kWaitWave = 1, kWaitBlock = 2, kDone = 3 };

struct Fiber {
  void* sp = nullptr;
  unsigned char* stack = nullptr;  // lowest usable byte (a guard page sits below it)
  Idx3 tid{0, 0, 0};
  int lane = 0, wave = 0;
  int state = kReady;
  unsigned ncoll = 0;  // collectives this fiber has joined
  int op = 0;          // op of the pending collective
};

struct Wave {
  unsigned char buf[2][64 * kPayload];  // lane l's payload of b bytes at offset l * b
};

struct Launch {
  const std::function<void()>* fn = nullptr;
  Idx3 bdim{1, 1, 1}, gdim{1, 1, 1}, bidx{0, 0, 0};
  unsigned char* lds = nullptr;
  std::vector<Fiber> fibers;
  std::vector<Wave> waves;
};

Launch* g_launch = nullptr;
Fiber* g_cur = nullptr;
void* g_sched_sp = nullptr;
const void* g_sched_stack_bottom = nullptr;
size_t g_sched_stack_size = 0;
std::vector<unsigned char*> g_stacks;

[[noreturn]] void die(const char* msg) {
  fprintf(stderr, "emu: %s\n", msg);
  fflush(stderr);
  abort();
}

unsigned char* stack_for(size_t i) {
  while (g_stacks.size() <= i) {
    const size_t page = 4096;
    void* p = mmap(nullptr, kStack + page, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (p == MAP_FAILED) die("mmap of a fiber stack failed");
    mprotect(p, page, PROT_NONE);
    g_stacks.push_back(static_cast<unsigned char*>(p) + page);
  }
  return g_stacks[i];
}

// scheduler -> fiber
void resume(Fiber* f) {
  g_cur = f;
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, f->stack, kStack);
  emu_ctx_switch(&g_sched_sp, f->sp);
  __sanitizer_finish_switch_fiber(fake, &g_sched_stack_bottom, &g_sched_stack_size);
  g_cur = nullptr;
}

// fiber -> scheduler
void yield_to_sched(bool exiting) {
  Fiber* f = g_cur;
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(exiting ? nullptr : &fake, g_sched_stack_bottom, g_sched_stack_size);
  emu_ctx_switch(&f->sp, g_sched_sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
}

}  // namespace

extern "C" void emu_fiber_main() {
  __sanitizer_finish_switch_fiber(nullptr, &g_sched_stack_bottom, &g_sched_stack_size);
  (*g_launch->fn)();
  g_cur->state = kDone;
  yield_to_sched(true);
  die("a finished fiber was resumed");
}

const Idx3& thread_idx() { return g_cur->tid; }
const Idx3& block_idx() { return g_launch->bidx; }
const Idx3& block_dim() { return g_launch->bdim; }
const Idx3& grid_dim() { return g_launch->gdim; }
unsigned char* dyn_lds() { return g_launch->lds; }
int lane_id() { return g_cur->lane; }

const unsigned char* wave_exchange(int op, const void* payload, int bytes) {
  if (bytes > kPayload) die("collective payload too large");
  Fiber* f = g_cur;
  const int par = f->ncoll & 1;
  unsigned char* slot = g_launch->waves[f->wave].buf[par] + (size_t)f->lane * bytes;
  if (bytes) memcpy(slot, payload, bytes);
  f->op = op;
  f->state = kWaitWave;
  yield_to_sched(false);
  f->ncoll++;
  return g_launch->waves[f->wave].buf[par];
}

void block_barrier() {
  g_cur->state = kWaitBlock;
  yield_to_sched(false);
}

namespace {

void init_fiber(Fiber& f, size_t i) {
  f.stack = stack_for(i);
  // initial frame for emu_ctx_switch: six callee-saved registers, then the entry trampoline
  uintptr_t top = (reinterpret_cast<uintptr_t>(f.stack) + kStack) & ~uintptr_t(15);
  void** sp = reinterpret_cast<void**>(top);
  *--sp = reinterpret_cast<void*>(&emu_fiber_entry);
  for (int r = 0; r < 6; ++r) *--sp = nullptr;
  f.sp = sp;
  f.state = kReady;
  f.ncoll = 0;
}

void run_block(Launch& L) {
  const int nt = (int)(L.bdim.x * L.bdim.y * L.bdim.z);
  for (int t = 0; t < nt; ++t) init_fiber(L.fibers[t], t);
  const int nw = (nt + 63) / 64;
  for (;;) {
    for (int t = 0; t < nt; ++t)
      if (L.fibers[t].state == kReady) resume(&L.fibers[t]);
    bool released = false;
    for (int w = 0; w < nw; ++w) {
      int waiting = 0, lanes = 0, op = -1;
      unsigned gen = 0;
      bool mismatch = false;
      for (int t = 64 * w; t < nt && t < 64 * w + 64; ++t) {
        const Fiber& f = L.fibers[t];
        ++lanes;
        if (f.state != kWaitWave) continue;
        if (waiting == 0) {
          op = f.op;
          gen = f.ncoll;
        } else if (f.op != op || f.ncoll != gen) {
          mismatch = true;
        }
        ++waiting;
      }
      if (waiting == 0) continue;
      if (lanes != 64) die("a wavefront collective in a workgroup that is not a whole number of wavefronts");
      if (waiting != 64) die("wavefront collective reached by only some lanes (divergent control flow)");
      if (mismatch) die("lanes of one wavefront are in different collectives");
      for (int t = 64 * w; t < 64 * w + 64; ++t) L.fibers[t].state = kReady;
      released = true;
    }
    if (released) continue;
    int at_barrier = 0, done = 0;
    for (int t = 0; t < nt; ++t) {
      at_barrier += L.fibers[t].state == kWaitBlock;
      done += L.fibers[t].state == kDone;
    }
    if (done == nt) return;
    if (at_barrier + done == nt) {
      if (done) die("__syncthreads with some threads of the workgroup already returned");
      for (int t = 0; t < nt; ++t) L.fibers[t].state = kReady;
      continue;
    }
    die("deadlock: no fiber can proceed");
  }
}

int env_int(const char* name, int dflt) {
  const char* v = getenv(name);
  return v ? atoi(v) : dflt;
}

}  // namespace

void launch(const std::function<void()>& k, dim3 grid, dim3 block, size_t lds_bytes) {
  if (g_launch) die("nested launch");
  const int nt = (int)(block.x * block.y * block.z);
  if (nt <= 0 || nt > 1024) die("invalid workgroup size");
  if (lds_bytes > 160 * 1024) die("dynamic LDS beyond 160 KiB");
  Launch L;
  L.fn = &k;
  L.bdim = Idx3{block.x, block.y, block.z};
  L.gdim = Idx3{grid.x, grid.y, grid.z};
  L.fibers.resize(nt);
  L.waves.resize((nt + 63) / 64);
  for (int t = 0; t < nt; ++t) {
    Fiber& f = L.fibers[t];
    f.tid = Idx3{(unsigned)t % block.x, ((unsigned)t / block.x) % block.y, (unsigned)t / (block.x * block.y)};
    f.lane = t & 63;
    f.wave = t >> 6;
  }
  g_launch = &L;
  for (unsigned z = 0; z < grid.z; ++z)
    for (unsigned y = 0; y < grid.y; ++y)
      for (unsigned x = 0; x < grid.x; ++x) {
        L.bidx = Idx3{x, y, z};
        // a fresh allocation of exactly the launch's dynamic LDS (filled with a NaN pattern: the
        // hardware does not clear LDS between workgroups either)
        L.lds = static_cast<unsigned char*>(malloc(lds_bytes ? lds_bytes : 1));
        memset(L.lds, env_int("EMU_LDS_FILL", 0xff), lds_bytes ? lds_bytes : 1);
        {  // optional NaN window over [EMU_LDS_NAN_LO, EMU_LDS_NAN_HI) bytes (localising reads of unwritten LDS)
          const int lo = env_int("EMU_LDS_NAN_LO", 0), hi = env_int("EMU_LDS_NAN_HI", 0);
          for (int i = lo; i < hi && i < (int)lds_bytes; ++i) L.lds[i] = 0xff;
        }
        run_block(L);
        free(L.lds);
        L.lds = nullptr;
      }
  g_launch = nullptr;
}

}  // namespace emu

// ------------------------------------------------------------------ HIP host API on host memory
struct ihipStream_t {
  int dummy;
};
struct ihipEvent_t {
  int dummy;
};

hipError_t hipMalloc(void** p, size_t bytes) {
  *p = malloc(bytes ? bytes : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) { return hipMalloc(p, bytes); }
hipError_t hipMemcpy(void* d, const void* s, size_t b, hipMemcpyKind) {
  if (b) memcpy(d, s, b);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t b, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(d, s, b, k);
}
hipError_t hipMemsetAsync(void* d, int v, size_t b, hipStream_t) {
  if (b) memset(d, v, b);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
  if (a == hipDeviceAttributeMaxSharedMemoryPerBlock) *v = 160 * 1024;
  else if (a == hipDeviceAttributeMultiprocessorCount) *v = getenv("EMU_CUS") ? atoi(getenv("EMU_CUS")) : 2;
  else return hipErrorInvalidValue;
  return hipSuccess;
}
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* nb, const void*, int threads, size_t lds) {
  int by_lds = lds ? (int)((160 * 1024) / lds) : 8;
  int by_thr = 2048 / (threads > 0 ? threads : 64);
  int v = by_lds < by_thr ? by_lds : by_thr;
  *nb = v < 1 ? 1 : (v > 8 ? 8 : v);
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* st, unsigned) {
  *st = new ihipStream_t{0};
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t st) {
  delete st;
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new ihipEvent_t{0};
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 1.0f;
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "hipSuccess" : "emulated HIP error"; }

// exact-size buffers for the test driver (ctypes)
extern "C" __attribute__((visibility("default"))) void* emu_alloc(size_t bytes) { return malloc(bytes ? bytes : 1); }
extern "C" __attribute__((visibility("default"))) void emu_free(void* p) { free(p); }
