// pool_kernels.hip - launches of the persistent pool kernels (batched and
// served shots) and their queue kernels. The device code is in
// step_device.h.
#include <hip/hip_runtime.h>

#include "step_device.h"

namespace fdtd2d {
namespace {

__global__ void __launch_bounds__(64) pool_tick(PoolQueue *q, int nk, int dg) {
  int *t = &q->ticket[0][0][0];
  for (int i = threadIdx.x; i < nk * kPoolLists * 8; i += blockDim.x) t[i] = 0;
  if (threadIdx.x == 0) q->g += dg;
}

struct PoolLists {
  const int4 *items[kPoolLists];
  int count[kPoolLists];
};

__global__ void __launch_bounds__(64) pool_set(PoolQueue *q, PoolLists l, int g) {
  if (threadIdx.x < kPoolLists) {
    q->items[threadIdx.x] = l.items[threadIdx.x];
    q->count[threadIdx.x] = l.count[threadIdx.x];
  }
  if (threadIdx.x == 0 && g >= 0) q->g = g;
}

}  // namespace

hipError_t launch_pool(const StepArgs *descs, PoolQueue *q, int k, int tp, int cls, int T, int C, int D,
                       int medium_kind, int grid, hipStream_t stream) {
  if (grid <= 0 || k < 0 || k >= kMaxChunk || cls < 0 || cls > 3 || (C != 1 && C != 2) || (D != 1 && D != 2))
    return hipErrorInvalidValue;
  if (cls == 1) return launch_pool_d<1, 1, 1>(descs, q, k, tp, T, medium_kind, grid, stream);
  if (cls == 3) return launch_pool_d<1, 1, 3>(descs, q, k, tp, T, medium_kind, grid, stream);
  if (C == 2) return launch_pool_c2(descs, q, k, tp, cls, T, D, medium_kind, grid, stream);
  if (cls == 2) {
    if (D == 2) return launch_pool_d<1, 2, 2>(descs, q, k, tp, T, medium_kind, grid, stream);
    return launch_pool_d<1, 1, 2>(descs, q, k, tp, T, medium_kind, grid, stream);
  }
  if (D == 2) return launch_pool_d<1, 2, 0>(descs, q, k, tp, T, medium_kind, grid, stream);
  return launch_pool_d<1, 1, 0>(descs, q, k, tp, T, medium_kind, grid, stream);
}

hipError_t launch_pool_set(PoolQueue *q, const int4 *const items[kPoolLists], const int count[kPoolLists], int g,
                           hipStream_t stream) {
  PoolLists l;
  for (int c = 0; c < kPoolLists; ++c) {
    l.items[c] = items[c];
    l.count[c] = count[c];
  }
  hipLaunchKernelGGL(pool_set, dim3(1), dim3(64), 0, stream, q, l, g);
  return hipGetLastError();
}

hipError_t launch_pool_tick(PoolQueue *q, int nk, int dg, hipStream_t stream) {
  if (nk < 0 || nk > kMaxChunk) return hipErrorInvalidValue;
  hipLaunchKernelGGL(pool_tick, dim3(1), dim3(64), 0, stream, q, nk, dg);
  return hipGetLastError();
}

}  // namespace fdtd2d
