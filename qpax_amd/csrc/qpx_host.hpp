// Host-side pieces shared by the translation units of libqpx: error reporting, workspace sizing
// and the launchers of the kernel variants (each defined beside its kernel).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/qpx.h"
#include "qpx_args.hpp"

namespace qpx {
namespace host {

// Records the message for qpx_last_error() and returns `st`.
qpx_status fail(qpx_status st, const char* fmt, ...);
void clear_error();
const char* last_error();

#define QPX_HIP(call)                                                                                \
  do {                                                                                               \
    hipError_t e_ = (call);                                                                          \
    if (e_ != hipSuccess) return ::qpx::host::fail(QPX_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// Head of every workspace: counters (word 0: plain-mode active count) and work-queue heads.
constexpr size_t kCtlBytes = 4096;
constexpr int kMaxQueues = 512;  // queue heads at words 1..kMaxQueues of the control block

// float64 words of workspace one problem of size (n, m_eq, m_ineq) occupies (any kernel).
int64_t stride_words(int n, int me, int mi);
int max_lds_per_block();
int cu_count();
// Polls a split relaxation waits for its solve: `dflt`, or the environment variable
// QPX_SPLIT_SPIN when set (a test lowers it to 0 to make every split wait give up).
int split_spin_limit(int dflt);

// Wavefront-per-QP kernel instance (NMAX, MIC) for a problem, or false if it does not fit one.
struct WaveVariant {
  int nmax, mic;
};
bool wave_variant(int n, int me, int mi, int ts, WaveVariant* v);

// Workgroup-per-QP matrix-core kernel instance (NTI 16-row tiles per side of the augmented KKT
// matrix, W wavefronts), or false if the problem does not fit one.  `w` = 0 picks the default.
struct GroupVariant {
  int nti, w;
};
bool group_variant(int n, int me, int mi, int ts, int w, GroupVariant* v);
int group_lds_bytes(int n, int me, int mi, int ts, GroupVariant v);

}  // namespace host

// Launchers of the workgroup-per-QP kernels (any size), NT in {64, 128, 256}.
template <int NT, typename TS>
hipError_t launch_generic_solve(const Args<TS>& a, unsigned grid, int lds, hipStream_t st, int relax);
template <int NT, typename TS>
hipError_t launch_generic_bwd(const Args<TS>& a, unsigned grid, int lds, hipStream_t st);
template <int NT, typename TS>
int generic_occupancy(int lds);  // resident workgroups per CU

namespace group {
template <typename TS, int NTI, int W>
hipError_t launch_group_solve(const Args<TS>& a, unsigned grid, int lds, hipStream_t st, int relax);
template <typename TS, int NTI, int W>
hipError_t launch_group_bwd(const Args<TS>& a, unsigned grid, int lds, hipStream_t st);
template <typename TS, int NTI, int W>
int group_occupancy(int lds);

// Runtime selection among the compiled (NTI, W) instances: NTI in 4..10, W in {4, 8}.
#define QPX_GROUP_SWITCH(V, CALL)                                            \
  switch (V.nti * 16 + V.w) {                                                \
    case 4 * 16 + 4: return CALL(4, 4);   case 4 * 16 + 8: return CALL(4, 8);   \
    case 5 * 16 + 4: return CALL(5, 4);   case 5 * 16 + 8: return CALL(5, 8);   \
    case 6 * 16 + 4: return CALL(6, 4);   case 6 * 16 + 8: return CALL(6, 8);   \
    case 7 * 16 + 4: return CALL(7, 4);   case 7 * 16 + 8: return CALL(7, 8);   \
    case 8 * 16 + 4: return CALL(8, 4);   case 8 * 16 + 8: return CALL(8, 8);   \
    case 9 * 16 + 4: return CALL(9, 4);   case 9 * 16 + 8: return CALL(9, 8);   \
    case 10 * 16 + 4: return CALL(10, 4); default: return CALL(10, 8);          \
  }

template <typename TS>
inline hipError_t launch_solve(host::GroupVariant v, const Args<TS>& a, unsigned grid, int lds, hipStream_t st, int relax) {
#define QPX_C(N, W) launch_group_solve<TS, N, W>(a, grid, lds, st, relax)
  QPX_GROUP_SWITCH(v, QPX_C)
#undef QPX_C
}
template <typename TS>
inline hipError_t launch_bwd(host::GroupVariant v, const Args<TS>& a, unsigned grid, int lds, hipStream_t st) {
#define QPX_C(N, W) launch_group_bwd<TS, N, W>(a, grid, lds, st)
  QPX_GROUP_SWITCH(v, QPX_C)
#undef QPX_C
}
template <typename TS>
inline int occupancy(host::GroupVariant v, int lds) {
#define QPX_C(N, W) group_occupancy<TS, N, W>(lds)
  QPX_GROUP_SWITCH(v, QPX_C)
#undef QPX_C
}
}  // namespace group

namespace wave {
template <typename TS, int NMAX, int MIC>
int wave_occupancy(int wpb, int lds_bytes);  // resident workgroups of wpb wavefronts per CU
template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_solve(const Args<TS>& a, unsigned grid, int wpb, hipStream_t st, int relax);
template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_bwd(const Args<TS>& a, unsigned grid, int wpb, hipStream_t st);

// Runtime selection among the compiled (NMAX, MIC) instances.
#define QPX_WAVE_SWITCH(V, CALL)                   \
  if (V.nmax == 32 && V.mic == 1) return CALL(32, 1); \
  if (V.nmax == 32 && V.mic == 2) return CALL(32, 2); \
  if (V.nmax == 64 && V.mic == 1) return CALL(64, 1); \
  if (V.nmax == 64 && V.mic == 2) return CALL(64, 2); \
  return CALL(64, 4);

template <typename TS>
inline hipError_t launch_solve(host::WaveVariant v, const Args<TS>& a, unsigned grid, int wpb, hipStream_t st, int relax) {
#define QPX_C(N, M) launch_wave_solve<TS, N, M>(a, grid, wpb, st, relax)
  QPX_WAVE_SWITCH(v, QPX_C)
#undef QPX_C
}
template <typename TS>
inline hipError_t launch_bwd(host::WaveVariant v, const Args<TS>& a, unsigned grid, int wpb, hipStream_t st) {
#define QPX_C(N, M) launch_wave_bwd<TS, N, M>(a, grid, wpb, st)
  QPX_WAVE_SWITCH(v, QPX_C)
#undef QPX_C
}
template <typename TS>
inline int occupancy(host::WaveVariant v, int wpb, int lds_bytes) {
#define QPX_C(N, M) wave_occupancy<TS, N, M>(wpb, lds_bytes)
  QPX_WAVE_SWITCH(v, QPX_C)
#undef QPX_C
}
}  // namespace wave

}  // namespace qpx
