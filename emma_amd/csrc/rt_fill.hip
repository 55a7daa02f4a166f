// rt_fill.hip -- the work either side of a level step that couples the
// radiation state to the hydro state and places the test source: FillRad's
// per-cell block (src_utils.c:692-849) and RadSolver's injection back into
// the hydro fields (rad_utils.c:1941-1951, 1993-2004).  FP64 throughout with
// the reference's operation order (this file is compiled without
// contraction like the rest of the library).
#include <hip/hip_runtime.h>

#include "rt_fill.h"

namespace emma {

template <typename R>
__global__ __launch_bounds__(256) void fill_level_kernel(FillArgs<R> f) {
  const int cell = blockIdx.x * 256 + (int)threadIdx.x;
  if (cell < f.ncell) fill_cell<R>(f, cell);
}

__global__ __launch_bounds__(256) void rad_to_hydro_kernel(HydroArgs h) {
  const int i = blockIdx.x * 256 + (int)threadIdx.x;
  if (i < h.ncell) to_hydro_cell(h, i);
}

template <typename R>
hipError_t launch_fill_level(const FillArgs<R>& f, hipStream_t s) {
  if (f.ncell == 0) return hipSuccess;
  hipLaunchKernelGGL((fill_level_kernel<R>), dim3((f.ncell + 255) / 256), dim3(256), 0, s, f);
  return hipGetLastError();
}
template hipError_t launch_fill_level<double>(const FillArgs<double>&, hipStream_t);
template hipError_t launch_fill_level<float>(const FillArgs<float>&, hipStream_t);

hipError_t launch_rad_to_hydro(const HydroArgs& h, hipStream_t s) {
  if (h.ncell == 0) return hipSuccess;
  hipLaunchKernelGGL(rad_to_hydro_kernel, dim3((h.ncell + 255) / 256), dim3(256), 0, s, h);
  return hipGetLastError();
}

}  // namespace emma
