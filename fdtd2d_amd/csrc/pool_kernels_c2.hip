// pool_kernels_c2.hip - the pool kernels that carry interior tiles (classes
// 0 and 2) at two columns per lane (see pool_kernels.hip). The device code is
// in step_device.h.
#include <hip/hip_runtime.h>

#include "step_device.h"

namespace fdtd2d {

hipError_t launch_pool_c2(const StepArgs *descs, PoolQueue *q, int k, int tp, int cls, int T, int D,
                          int medium_kind, int grid, hipStream_t stream) {
  if (cls == 2) {
    if (D == 2) return launch_pool_d<2, 2, 2>(descs, q, k, tp, T, medium_kind, grid, stream);
    return launch_pool_d<2, 1, 2>(descs, q, k, tp, T, medium_kind, grid, stream);
  }
  if (cls != 0) return hipErrorInvalidValue;
  if (D == 2) return launch_pool_d<2, 2, 0>(descs, q, k, tp, T, medium_kind, grid, stream);
  return launch_pool_d<2, 1, 0>(descs, q, k, tp, T, medium_kind, grid, stream);
}

}  // namespace fdtd2d
