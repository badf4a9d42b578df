// Step kernels of one level capacity (lbm_levels.h), built once per
// capacity with -DLBM_LEVEL_CAP=16, 32, 48 or 64.
#include "lbm_levels.h"

#ifndef LBM_LEVEL_CAP
#error "build with -DLBM_LEVEL_CAP=16|32|48|64"
#endif

namespace lbm {

template <class T, int CAP>
hipError_t launch_step_cap(const StepArgs<T> &a, dim3 grid, hipStream_t st) {
  hipLaunchKernelGGL((k_step<T, CAP, false, false>), grid, dim3(64), 0, st, a);
  return hipGetLastError();
}

template <class T, int CAP>
hipError_t launch_step_adj_cap(const StepAdjArgs<T> &a, dim3 grid, hipStream_t st) {
  hipLaunchKernelGGL((k_step_adj<T, CAP, false>), grid, dim3(64), 0, st, a);
  return hipGetLastError();
}

template hipError_t launch_step_cap<double, LBM_LEVEL_CAP>(const StepArgs<double> &, dim3, hipStream_t);
template hipError_t launch_step_cap<float, LBM_LEVEL_CAP>(const StepArgs<float> &, dim3, hipStream_t);
template hipError_t launch_step_adj_cap<double, LBM_LEVEL_CAP>(const StepAdjArgs<double> &, dim3,
                                                              hipStream_t);
template hipError_t launch_step_adj_cap<float, LBM_LEVEL_CAP>(const StepAdjArgs<float> &, dim3,
                                                             hipStream_t);

}  // namespace lbm
