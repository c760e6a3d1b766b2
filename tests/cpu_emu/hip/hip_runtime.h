// Host build of libqpx's HIP sources for the CPU emulator used by the bounds tests.
//
// The library's kernels are compiled for x86-64 against this header instead of ROCm's: every
// thread of a launched workgroup runs as a fiber on one host thread, and the wavefront-wide
// operations (readlane, DPP moves, shuffles, ballots, the float64 MFMAs) and the barriers are
// collectives of the fibers of one wavefront or workgroup.  Device memory is host memory
// allocated with its exact size, and each workgroup's dynamic LDS is a fresh allocation of
// exactly the launch's byte count, so an AddressSanitizer build reports any access outside a
// buffer, the LDS image included.  Test infrastructure only: nothing in the product loads it.
#pragma once

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>
#include <stdio.h>

#include <cmath>
#include <functional>
#include <type_traits>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))

// the kernels' register-pinning asm and their dynamic-LDS declaration, in host form
#define QPX_PIN_V(x) asm volatile("" : "+r"(x))
#define QPX_PIN_S(x) asm volatile("" : "+r"(x))
#define QPX_DYN_LDS(name) unsigned char* const name = ::emu::dyn_lds()
#define QPX_GLOBAL_AS
#define QPX_LDS_AS
#define QPX_ASSUME_LDS(p) ((void)0)

struct alignas(16) double2 {
  double x, y;
};
struct alignas(8) float2 {
  float x, y;
};
struct alignas(16) float4 {
  float x, y, z, w;
};

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned x_ = 1, unsigned y_ = 1, unsigned z_ = 1) : x(x_), y(y_), z(z_) {}
};

typedef enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 } hipError_t;
typedef enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
} hipMemcpyKind;
typedef enum hipDeviceAttribute_t {
  hipDeviceAttributeMaxSharedMemoryPerBlock = 1,
  hipDeviceAttributeMultiprocessorCount = 2
} hipDeviceAttribute_t;
struct ihipStream_t;
struct ihipEvent_t;
typedef ihipStream_t* hipStream_t;
typedef ihipEvent_t* hipEvent_t;
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u
#define __HIP_MEMORY_SCOPE_AGENT 3

hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t st);
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t st);
hipError_t hipDeviceSynchronize();
hipError_t hipGetDevice(int* dev);
hipError_t hipSetDevice(int dev);
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t attr, int dev);
hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* nb, const void* fn, int threads, size_t lds);
hipError_t hipStreamCreateWithFlags(hipStream_t* st, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t st);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamWaitEvent(hipStream_t st, hipEvent_t e, unsigned flags);
hipError_t hipGetLastError();
#define HIP_SYMBOL(X) X
template <typename T>
inline hipError_t hipGetSymbolAddress(void** p, const T& symbol) {
  *p = (void*)&symbol;
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t e);

namespace emu {

struct Idx3 {
  unsigned x, y, z;
};
const Idx3& thread_idx();
const Idx3& block_idx();
const Idx3& block_dim();
const Idx3& grid_dim();
unsigned char* dyn_lds();
// Runs `k` once per thread of every workgroup of the grid (workgroups one after another).
void launch(const std::function<void()>& k, dim3 grid, dim3 block, size_t lds_bytes);

// Wavefront collectives: every lane of the calling wavefront deposits `bytes` of payload and
// waits until all 64 lanes have deposited; the return value is the 64 payloads (lane-major,
// `bytes` apart), valid until this lane's next collective.
const unsigned char* wave_exchange(int op, const void* payload, int bytes);
void block_barrier();
int lane_id();

enum Op { kReadlane = 1, kReadfirst, kDpp, kShfl, kBallot, kMfma16, kMfma4, kWaveBarrier, kAtomicLoad };

typedef double d4 __attribute__((ext_vector_type(4)));

template <typename T>
inline T bcast_from(int op, T v, int src) {
  const unsigned char* all = wave_exchange(op, &v, (int)sizeof(T));
  T r;
  memcpy(&r, all + sizeof(T) * (size_t)(src & 63), sizeof(T));
  return r;
}

inline int readlane(int v, int lane) { return bcast_from(kReadlane, v, lane); }
inline int readfirstlane(int v) { return bcast_from(kReadfirst, v, 0); }

// DPP lane moves of the controls the kernels use (quad_perm, row_mirror, row_half_mirror); all
// row and bank masks enabled.
inline int dpp_src(int lane, int ctrl) {
  if (ctrl >= 0 && ctrl <= 0xff) return (lane & ~3) | ((ctrl >> (2 * (lane & 3))) & 3);
  if (ctrl == 0x140) return (lane & ~15) | (15 - (lane & 15));
  if (ctrl == 0x141) return (lane & ~7) | (7 - (lane & 7));
  __builtin_trap();
}
inline int mov_dpp(int v, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
  (void)bound_ctrl;
  if (row_mask != 0xf || bank_mask != 0xf) __builtin_trap();
  return bcast_from(kDpp, v, dpp_src(lane_id(), ctrl));
}

// v_mfma_f64_16x16x4f64: A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in register
// i / 4 of lane j + 16 (i % 4).
inline d4 mfma16(double a, double b, d4 acc) {
  struct P {
    double a, b;
  } p{a, b};
  const P* all = reinterpret_cast<const P*>(wave_exchange(kMfma16, &p, (int)sizeof(P)));
  const int l = lane_id(), j = l & 15;
  d4 out = acc;
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * r + (l >> 4);
    double s = acc[r];
    for (int k = 0; k < 4; ++k) s += all[i + 16 * k].a * all[j + 16 * k].b;
    out[r] = s;
  }
  return out;
}
// v_mfma_f64_4x4x4f64 (four blocks): block b = (lane >> 2) & 3 holds A[i][k] in lane
// 16k + 4b + i, B[k][j] in lane 16k + 4b + j and D[i][j] in lane 16i + 4b + j.
inline double mfma4(double a, double b, double acc) {
  struct P {
    double a, b;
  } p{a, b};
  const P* all = reinterpret_cast<const P*>(wave_exchange(kMfma4, &p, (int)sizeof(P)));
  const int l = lane_id(), i = l >> 4, bk = (l >> 2) & 3, j = l & 3;
  double s = acc;
  for (int k = 0; k < 4; ++k) s += all[16 * k + 4 * bk + i].a * all[16 * k + 4 * bk + j].b;
  return s;
}

inline uint64_t ballot(int pred) {
  const unsigned char* all = wave_exchange(kBallot, &pred, (int)sizeof(int));
  uint64_t m = 0;
  for (int l = 0; l < 64; ++l) {
    int v;
    memcpy(&v, all + sizeof(int) * l, sizeof(int));
    if (v) m |= 1ull << l;
  }
  return m;
}

inline void wave_barrier() { (void)wave_exchange(kWaveBarrier, nullptr, 0); }

}  // namespace emu

#define threadIdx (::emu::thread_idx())
#define blockIdx (::emu::block_idx())
#define blockDim (::emu::block_dim())
#define gridDim (::emu::grid_dim())

#define __builtin_amdgcn_readlane(v, l) ::emu::readlane((v), (l))
#define __builtin_amdgcn_readfirstlane(v) ::emu::readfirstlane((v))
#define __builtin_amdgcn_mov_dpp(v, c, rm, bm, bc) ::emu::mov_dpp((v), (c), (rm), (bm), (bc))
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) ::emu::mfma16((a), (b), (c))
#define __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, x, y, z) ::emu::mfma4((a), (b), (c))
#define __builtin_amdgcn_rcp(d) (1.0 / (d))
// v_rsq_f64 is an approximation that callers refine: the seed here is off by 2^-20 so the
// refinement is exercised
#define __builtin_amdgcn_rsq(d) ((1.0 + 0x1p-20) / std::sqrt((double)(d)))
#define __builtin_amdgcn_sched_barrier(m) ((void)0)
#define __builtin_amdgcn_wave_barrier() ::emu::wave_barrier()
#define __builtin_amdgcn_s_sleep(n) ((void)0)

inline void __syncthreads() { ::emu::block_barrier(); }
// A fence orders the calling wavefront's accesses; the lanes of a wavefront execute in lockstep on
// the hardware, so for the fibers of one wavefront it is a wavefront barrier.
inline void __threadfence() { ::emu::wave_barrier(); }
inline void __threadfence_block() { ::emu::wave_barrier(); }

inline double __shfl(double v, int src, int width = 64) {
  (void)width;
  return ::emu::bcast_from(::emu::kShfl, v, src);
}
inline int __shfl(int v, int src, int width = 64) {
  (void)width;
  return ::emu::bcast_from(::emu::kShfl, v, src);
}
inline float __shfl(float v, int src, int width = 64) {
  (void)width;
  return ::emu::bcast_from(::emu::kShfl, v, src);
}
template <typename T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  (void)width;
  return ::emu::bcast_from(::emu::kShfl, v, ::emu::lane_id() ^ mask);
}
inline uint64_t __ballot(int pred) { return ::emu::ballot(pred); }

inline int __double2loint(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return (int)(uint32_t)u;
}
inline int __double2hiint(double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  return (int)(uint32_t)(u >> 32);
}
inline double __hiloint2double(int hi, int lo) {
  const uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
  double d;
  memcpy(&d, &u, 8);
  return d;
}

template <typename T>
inline T atomicAdd(T* p, T v) {
  const T old = *p;
  *p = old + v;
  return old;
}
namespace emu {
// A wavefront's atomic load is one access: every lane sees the value the first lane read.
template <typename T>
inline T atomic_load(T* p) {
  const T v = lane_id() == 0 ? *p : T();
  return bcast_from(kAtomicLoad, v, 0);
}
}  // namespace emu
#define __hip_atomic_load(p, order, scope) ::emu::atomic_load(p)
#define __hip_atomic_store(p, v, order, scope) ((void)(*(p) = (v)))

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }

using std::isinf;
using std::isnan;

#define hipLaunchKernelGGL(K, G, B, S, ST, ...) \
  ((void)(ST), ::emu::launch([&]() { K(__VA_ARGS__); }, dim3(G), dim3(B), (size_t)(S)))
