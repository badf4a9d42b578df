// Launchers of the spectral step and its adjoint for level counts other
// than 20 and 40.  The kernels keep a column's levels in registers up to a
// capacity of 16, 32, 48 or 64 and take the level count at run time; each
// capacity is compiled from lbm_levels.hip as a translation unit of its own
// (-DLBM_LEVEL_CAP=16|32|48|64), so that the library's objects build in
// parallel.
#pragma once
#include "lbm_adjoint.h"

namespace lbm {

template <class T, int CAP>
hipError_t launch_step_cap(const StepArgs<T> &a, dim3 grid, hipStream_t st);
template <class T, int CAP>
hipError_t launch_step_adj_cap(const StepAdjArgs<T> &a, dim3 grid, hipStream_t st);

// k_step / k_step_adj on (grid, 64 threads) for a.L levels (1 .. 64, not 20
// or 40); returns the launch status.
template <class T>
inline hipError_t launch_step_wide(const StepArgs<T> &a, dim3 grid, hipStream_t st) {
  if (a.L < 1) return hipErrorInvalidValue;
  if (a.L <= 16) return launch_step_cap<T, 16>(a, grid, st);
  if (a.L <= 32) return launch_step_cap<T, 32>(a, grid, st);
  if (a.L <= 48) return launch_step_cap<T, 48>(a, grid, st);
  if (a.L <= 64) return launch_step_cap<T, 64>(a, grid, st);
  return hipErrorInvalidValue;
}
template <class T>
inline hipError_t launch_step_adj_wide(const StepAdjArgs<T> &a, dim3 grid, hipStream_t st) {
  if (a.L < 1) return hipErrorInvalidValue;
  if (a.L <= 16) return launch_step_adj_cap<T, 16>(a, grid, st);
  if (a.L <= 32) return launch_step_adj_cap<T, 32>(a, grid, st);
  if (a.L <= 48) return launch_step_adj_cap<T, 48>(a, grid, st);
  if (a.L <= 64) return launch_step_adj_cap<T, 64>(a, grid, st);
  return hipErrorInvalidValue;
}

}  // namespace lbm
