// Host-side launchers of the wavefront-per-QP kernels.  Each (storage type, NMAX, MIC)
// instantiation is compiled in its own object from qpx_wave_inst.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "qpx_args.hpp"

namespace qpx {
namespace wave {

struct Lds {
  int off_G, off_A, off_L, off_c, off_v, bytes;
};
__host__ __device__ inline int al16(int v) { return (v + 15) & ~15; }
// LDS carve-up of one wavefront's problem: G and A in input precision, the packed factor (at
// least 256 doubles: it doubles as the G^T D G tile staging area), two 64-entry broadcast
// buffers, one m_ineq-entry broadcast buffer.
__host__ __device__ inline Lds lds_layout(int n, int me, int mi, int ts) {
  const int m = n + me;
  Lds w;
  int o = 0;
  w.off_G = o; o = al16(o + mi * n * ts);
  w.off_A = o; o = al16(o + me * n * ts);
  w.off_L = o; o = al16(o + 8 * (m * (m + 1) / 2 > 256 ? m * (m + 1) / 2 : 256));
  w.off_c = o; o = al16(o + 8 * 128);
  w.off_v = o; o = al16(o + 8 * (mi > 64 ? mi : 64));
  w.bytes = o;
  return w;
}

template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_solve(const Args<TS>& a, unsigned grid, int lds, hipStream_t st, int relax);
template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_bwd(const Args<TS>& a, unsigned grid, int lds, hipStream_t st);

}  // namespace wave
}  // namespace qpx
