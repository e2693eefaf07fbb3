// fused_kernels_c1.hip - the fused single-grid launch at one column per lane
// (see fused_kernels.hip). The device code is in step_device.h.
#include <hip/hip_runtime.h>

#include "step_device.h"

namespace fdtd2d {

hipError_t launch_step_fused_c1(const StepArgs &a, int T, int D, int src, int n0, hipStream_t stream) {
  switch (D) {
    case 1: return launch_fused_cd<1, 1>(a, T, src, n0, stream);
    case 2: return launch_fused_cd<1, 2>(a, T, src, n0, stream);
    case 4: return launch_fused_cd<1, 4>(a, T, src, n0, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace fdtd2d
