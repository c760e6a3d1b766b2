// One instantiation of the wavefront-per-QP kernels, selected at compile time with
// -DQPX_TS=<float|double> -DQPX_NMAX=<32|64> -DQPX_MIC=<1|2|4>.
#include "qpx_host.hpp"
#include "qpx_wave.hpp"

namespace qpx {
namespace wave {

// `grid` workgroups of `wpb` wavefronts; a.lds_per_wave bytes of LDS per wavefront.
template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_solve(const Args<TS>& a, unsigned grid, int wpb, hipStream_t st, int relax) {
  if (!a.rg && !a.qhead && !a.order && relax && wpb == 1 && (int64_t)grid == a.B && a.B > 0) {
    // uniform batch: solves and relaxations as separate work items, tickets from control word 1
    // the last 8 problems per CU (two per SIMD) run their relaxations as separate items; the rest
    // relax in their solve's item and skip the second load
    Args<TS> b = a;
    b.qhead = a.counter + 1;
    const int64_t split = 8 * (int64_t)qpx::host::cu_count();
    b.split_fused = a.B > split ? a.B - split : 0;
    b.split_spin = qpx::host::split_spin_limit(kSplitSpin);
    const int64_t flag_word = split_state_off(a.n, a.me) + kSplitStateWords;
    hipLaunchKernelGGL((k_wave_split_clear<TS>), dim3((unsigned)((a.B + 255) / 256)), dim3(256), 0, st, b, flag_word);
    hipLaunchKernelGGL((k_wave_solve<TS, NMAX, MIC, true>), dim3((unsigned)(2 * a.B - b.split_fused)), dim3(64),
                       a.lds_per_wave, st, b, 2);
    return hipGetLastError();
  }
  if (a.rg)
    hipLaunchKernelGGL((k_wave_solve<TS, NMAX, MIC, false>), dim3(grid), dim3(64 * wpb), a.lds_per_wave * wpb, st, a,
                       relax);
  else
    hipLaunchKernelGGL((k_wave_solve<TS, NMAX, MIC, true>), dim3(grid), dim3(64 * wpb), a.lds_per_wave * wpb, st, a,
                       relax);
  return hipGetLastError();
}

template <typename TS, int NMAX, int MIC>
hipError_t launch_wave_bwd(const Args<TS>& a, unsigned grid, int wpb, hipStream_t st) {
  if (a.rg)
    hipLaunchKernelGGL((k_wave_bwd<TS, NMAX, MIC, false>), dim3(grid), dim3(64 * wpb), a.lds_per_wave * wpb, st, a);
  else
    hipLaunchKernelGGL((k_wave_bwd<TS, NMAX, MIC, true>), dim3(grid), dim3(64 * wpb), a.lds_per_wave * wpb, st, a);
  return hipGetLastError();
}

// Resident workgroups of `wpb` wavefronts per CU (forward kernel as ragged batches run it, the
// longer-running of the two).
template <typename TS, int NMAX, int MIC>
int wave_occupancy(int wpb, int lds_bytes) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_wave_solve<TS, NMAX, MIC, false>),
                                                   64 * wpb, (size_t)lds_bytes * wpb) != hipSuccess)
    nb = 0;
  return nb;
}

template hipError_t launch_wave_solve<QPX_TS, QPX_NMAX, QPX_MIC>(const Args<QPX_TS>&, unsigned, int, hipStream_t, int);
template int wave_occupancy<QPX_TS, QPX_NMAX, QPX_MIC>(int, int);
template hipError_t launch_wave_bwd<QPX_TS, QPX_NMAX, QPX_MIC>(const Args<QPX_TS>&, unsigned, int, hipStream_t);

}  // namespace wave
}  // namespace qpx
