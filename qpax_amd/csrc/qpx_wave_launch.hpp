// Host-side launchers of the wavefront-per-QP kernels.  Each (storage type, NMAX, MIC)
// instantiation is compiled in its own object from qpx_wave_inst.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "qpx_args.hpp"

namespace qpx {
namespace wave {

struct Lds {
  int off_G, off_L, off_P, bytes;
};
__host__ __device__ inline int al16(int v) { return (v + 15) & ~15; }
// Rows n..np-1 of the reduced KKT system are padding so that the equality block starts at a
// multiple of 4 (the factorisation eliminates 4 pivots per panel).
__host__ __device__ inline int padded_n(int n) { return (n + 3) & ~3; }
// LDS carve-up of one wavefront's problem: G in input precision, the packed unit-lower factor of
// the m x m reduced KKT matrix (m = padded_n(n) + m_eq), and a panel buffer of 64 x 8 doubles that
// doubles as the broadcast buffers (128 + 2 m_ineq doubles) outside the factorisation.
__host__ __device__ inline Lds lds_layout(int n, int me, int mi, int ts) {
  const int m = padded_n(n) + me;
  Lds w;
  int o = 0;
  w.off_G = o; o = al16(o + mi * n * ts);
  w.off_L = o; o = al16(o + 8 * (m * (m + 1) / 2));
  w.off_P = o; o = al16(o + 8 * (128 + 2 * mi > 512 ? 128 + 2 * mi : 512));
  w.bytes = o;
  return w;
}

// Split forward launches (uniform batches): per problem, the float64 state [x; y] (64 lanes), s
// and z (64 x MIC each, MIC <= 4) and a ready flag follow the m x m constant matrix in the
// problem's workspace slice.
__host__ __device__ inline int64_t split_state_off(int n, int me) {
  const int64_t m = padded_n(n) + me;
  return (m * m + 63) / 64 * 64;
}
constexpr int kSplitStateWords = 64 + 2 * 64 * 4;
__host__ __device__ inline int64_t split_words(int n, int me) { return split_state_off(n, me) + kSplitStateWords + 64; }

template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_solve(const Args<TS>& a, unsigned grid, int wpb, hipStream_t st, int relax);
template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_bwd(const Args<TS>& a, unsigned grid, int wpb, hipStream_t st);

}  // namespace wave
}  // namespace qpx
