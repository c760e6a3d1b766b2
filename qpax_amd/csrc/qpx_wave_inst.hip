// One instantiation of the wavefront-per-QP kernels, selected at compile time with
// -DQPX_TS=<float|double> -DQPX_NMAX=<32|64> -DQPX_MIC=<1|2|4>.
#include "qpx_wave.hpp"

namespace qpx {
namespace wave {

template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_solve(const Args<TS>& a, unsigned grid, int lds, hipStream_t st, int relax) {
  hipLaunchKernelGGL((k_wave_solve<TS, NMAX, MIC>), dim3(grid), dim3(64), lds, st, a, relax);
  return hipGetLastError();
}

template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_bwd(const Args<TS>& a, unsigned grid, int lds, hipStream_t st) {
  hipLaunchKernelGGL((k_wave_bwd<TS, NMAX, MIC>), dim3(grid), dim3(64), lds, st, a);
  return hipGetLastError();
}

template hipError_t launch_wave_solve<QPX_TS, QPX_NMAX, QPX_MIC>(const Args<QPX_TS>&, unsigned, int, hipStream_t, int);
template hipError_t launch_wave_bwd<QPX_TS, QPX_NMAX, QPX_MIC>(const Args<QPX_TS>&, unsigned, int, hipStream_t);

}  // namespace wave
}  // namespace qpx
