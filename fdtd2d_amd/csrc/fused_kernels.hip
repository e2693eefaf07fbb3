// fused_kernels.hip - the fused single-grid launch: one step kernel per launch
// over one list holding the CPML-class and the interior work-groups (the
// CPML-class ones first in each XCD's range). Two columns per lane here, one
// column per lane in fused_kernels_c1.hip (the two compile in parallel; four
// columns per lane have no fused form: their kernel would spill). The device code is in step_device.h.
#include <hip/hip_runtime.h>

#include "step_device.h"

namespace fdtd2d {

hipError_t launch_step_fused_c1(const StepArgs &a, int T, int D, int src, int n0, hipStream_t stream);

hipError_t launch_step_fused(const StepArgs &a, int T, int C, int D, int src, int n0, hipStream_t stream) {
  switch (C * 16 + D) {
    case 1 * 16 + 1:
    case 1 * 16 + 2:
    case 1 * 16 + 4: return launch_step_fused_c1(a, T, D, src, n0, stream);
    case 2 * 16 + 1: return launch_fused_cd<2, 1>(a, T, src, n0, stream);
    case 2 * 16 + 2: return launch_fused_cd<2, 2>(a, T, src, n0, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace fdtd2d
