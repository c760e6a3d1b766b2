// One instantiation of the workgroup-per-QP matrix-core kernels, selected at compile time with
// -DQPX_TS=<float|double> -DQPX_NTI=<16-row tiles per side> -DQPX_W=<wavefronts>.
#include "qpx_host.hpp"
#include "qpx_group.hpp"

namespace qpx {
namespace group {

template <typename TS, int NTI, int W>
hipError_t launch_group_solve(const Args<TS>& a, unsigned grid, int lds, hipStream_t st, int relax) {
  hipLaunchKernelGGL((k_group_solve<TS, NTI, W>), dim3(grid), dim3(64 * W), lds, st, a, relax);
  return hipGetLastError();
}

template <typename TS, int NTI, int W>
hipError_t launch_group_bwd(const Args<TS>& a, unsigned grid, int lds, hipStream_t st) {
  hipLaunchKernelGGL((k_group_bwd<TS, NTI, W>), dim3(grid), dim3(64 * W), lds, st, a);
  return hipGetLastError();
}

template <typename TS, int NTI, int W>
int group_occupancy(int lds) {
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_group_solve<TS, NTI, W>),
                                                   64 * W, (size_t)lds) != hipSuccess)
    nb = 0;
  return nb;
}

template hipError_t launch_group_solve<QPX_TS, QPX_NTI, QPX_W>(const Args<QPX_TS>&, unsigned, int, hipStream_t, int);
template hipError_t launch_group_bwd<QPX_TS, QPX_NTI, QPX_W>(const Args<QPX_TS>&, unsigned, int, hipStream_t);
template int group_occupancy<QPX_TS, QPX_NTI, QPX_W>(int);

}  // namespace group
}  // namespace qpx
