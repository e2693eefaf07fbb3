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

template <int T, int D, int MODE>
hipError_t launch_pool_t(const StepArgs *descs, PoolQueue *q, int k, int tp, int med, int grid, hipStream_t s) {
  if (med == kTable)
    hipLaunchKernelGGL((pool_kernel<T, 1, D, kTable, MODE>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, descs, q, k,
                       tp);
  else
    hipLaunchKernelGGL((pool_kernel<T, 1, D, kDensity, MODE>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, descs, q,
                       k, tp);
  return hipGetLastError();
}

template <int D, int MODE>
hipError_t launch_pool_d(const StepArgs *descs, PoolQueue *q, int k, int tp, int T, int med, int grid,
                         hipStream_t s) {
  switch (T) {
    case 1: return launch_pool_t<1, D, MODE>(descs, q, k, tp, med, grid, s);
    case 2: return launch_pool_t<2, D, MODE>(descs, q, k, tp, med, grid, s);
    case 3: return launch_pool_t<3, D, MODE>(descs, q, k, tp, med, grid, s);
    case 4: return launch_pool_t<4, D, MODE>(descs, q, k, tp, med, grid, s);
    case 5: return launch_pool_t<5, D, MODE>(descs, q, k, tp, med, grid, s);
    case 6: return launch_pool_t<6, D, MODE>(descs, q, k, tp, med, grid, s);
    case 7: return launch_pool_t<7, D, MODE>(descs, q, k, tp, med, grid, s);
    case 8: return launch_pool_t<8, D, MODE>(descs, q, k, tp, med, grid, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_pool(const StepArgs *descs, PoolQueue *q, int k, int tp, int cls, int T, int D, int medium_kind,
                       int grid, hipStream_t stream) {
  if (grid <= 0 || k < 0 || k >= kMaxChunk || cls < 0 || cls > 3) return hipErrorInvalidValue;
  if (cls == 1) return launch_pool_d<1, 1>(descs, q, k, tp, T, medium_kind, grid, stream);
  if (cls == 3) return launch_pool_d<1, 3>(descs, q, k, tp, T, medium_kind, grid, stream);
  if (cls == 2) {
    if (D == 2) return launch_pool_d<2, 2>(descs, q, k, tp, T, medium_kind, grid, stream);
    return launch_pool_d<1, 2>(descs, q, k, tp, T, medium_kind, grid, stream);
  }
  if (D == 2) return launch_pool_d<2, 0>(descs, q, k, tp, T, medium_kind, grid, stream);
  return launch_pool_d<1, 0>(descs, q, k, tp, T, medium_kind, grid, stream);
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
