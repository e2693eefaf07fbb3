// step_kernels.hip - single-grid launches of the step kernels (interior and
// CPML-class kernels on two streams, or one fused kernel: fused_kernels.hip)
// and the RC cutoff reduction kernel. The device code is in step_device.h.
#include <hip/hip_runtime.h>

#include "step_device.h"

namespace fdtd2d {
namespace {

// RC cutoff lines for the t steps of one launch, in step order (the edge level
// carries from step to step). Dev_FindMax keeps a value only when a later one
// is strictly greater, so the maximum is the first of the largest values in
// scan order (row 0, then rows jcut .. ny-3) and a NaN at row 0 stays; the
// scan (j from ny-3 down while |Ez| < edge) stops at the largest j >= 1 whose
// |Ez| is not below the edge, else at 0.
__global__ void __launch_bounds__(256) rc_kernel(const RcArgs r, int n0, int t) {
  __shared__ float s_v[256];
  __shared__ int s_i[256];
  const int e = blockIdx.x, tid = threadIdx.x;
  const int pm = r.map[2 * e], ps = r.map[2 * e + 1];
  const int ny = r.ny;
  const int ncand = 1 + max(0, (ny - 2) - r.jcut);
  float edge = r.edge[e];
  for (int s = 1; s <= t; ++s) {
    const int step = n0 + s;
    const long long slot = (step - 1) & (kMaxT - 1);
    const float *cm = r.cap + (slot * r.nprobe + pm) * ny;
    const float *cs = r.cap + (slot * r.nprobe + ps) * ny;
    // (value, candidate index): larger value wins, equal values the earlier
    float bv = 0.f;
    int bi = 0x7fffffff;
    for (int q = tid; q < ncand; q += blockDim.x) {
      const float v = q == 0 ? cm[0] : cm[r.jcut + q - 1];
      if (v != v) continue;  // NaN never replaces (row 0 handled below)
      if (bi == 0x7fffffff || v > bv || (v == bv && q < bi)) {
        bv = v;
        bi = q;
      }
    }
    s_v[tid] = bv;
    s_i[tid] = bi;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (tid < w) {
        const float ov = s_v[tid + w];
        const int oi = s_i[tid + w];
        if (oi != 0x7fffffff &&
            (s_i[tid] == 0x7fffffff || ov > s_v[tid] || (ov == s_v[tid] && oi < s_i[tid]))) {
          s_v[tid] = ov;
          s_i[tid] = oi;
        }
      }
      __syncthreads();
    }
    const float c0 = cm[0];
    const float temp = c0 != c0 ? c0 : s_v[0];
    __syncthreads();
    const float et = temp * r.level;
    if (et > edge) edge = et;
    int bj = 0;
    for (int j = 1 + tid; j <= ny - 3; j += blockDim.x)
      if (!(fabsf(cs[j]) < edge)) bj = max(bj, j);
    s_i[tid] = bj;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (tid < w) s_i[tid] = max(s_i[tid], s_i[tid + w]);
      __syncthreads();
    }
    if (tid == 0) r.spd[static_cast<long long>(step) * r.nrc + e] = s_i[0];
    __syncthreads();
  }
  if (tid == 0) r.edge[e] = edge;
}

template <int T, int C, int D, int MODE>
hipError_t launch_tcd(const StepArgs &a, int src, int n0, hipStream_t s) {
  const int grid = a.n_wg[(MODE == 1 || MODE == 3) ? 1 : 0];
  if (grid == 0) return hipSuccess;
  if (a.medium_kind == kTable)
    hipLaunchKernelGGL((step_kernel<T, C, D, kTable, MODE>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, a, src, n0);
  else
    hipLaunchKernelGGL((step_kernel<T, C, D, kDensity, MODE>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, a, src,
                       n0);
  return hipGetLastError();
}

template <int C, int D, int MODE>
hipError_t launch_cd(const StepArgs &a, int T, int src, int n0, hipStream_t s) {
  switch (T) {
    case 1: return launch_tcd<1, C, D, MODE>(a, src, n0, s);
    case 2: return launch_tcd<2, C, D, MODE>(a, src, n0, s);
    case 3: return launch_tcd<3, C, D, MODE>(a, src, n0, s);
    case 4: return launch_tcd<4, C, D, MODE>(a, src, n0, s);
    case 5: return launch_tcd<5, C, D, MODE>(a, src, n0, s);
    case 6: return launch_tcd<6, C, D, MODE>(a, src, n0, s);
    case 7: return launch_tcd<7, C, D, MODE>(a, src, n0, s);
    case 8: return launch_tcd<8, C, D, MODE>(a, src, n0, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_rc(const RcArgs &r, int n0, int t, hipStream_t stream) {
  if (r.nrc <= 0 || t <= 0) return hipSuccess;
  if (t > kMaxT) return hipErrorInvalidValue;
  hipLaunchKernelGGL(rc_kernel, dim3(r.nrc), dim3(256), 0, stream, r, n0, t);
  return hipGetLastError();
}

hipError_t launch_step(const StepArgs &a, int cls, int T, int C, int D, int src, int n0, hipStream_t stream) {
  // CPML-class work-groups always run one column per lane with the shallow
  // prefetch (their state already fills the register file); interior ones
  // take the requested shape.
  if (cls == 1) return launch_cd<1, 1, 1>(a, T, src, n0, stream);
  if (cls == 2) return launch_cd<1, 1, 3>(a, T, src, n0, stream);
  switch (C * 16 + D) {
    case 1 * 16 + 1: return launch_cd<1, 1, 0>(a, T, src, n0, stream);
    case 1 * 16 + 2: return launch_cd<1, 2, 0>(a, T, src, n0, stream);
    case 1 * 16 + 4: return launch_cd<1, 4, 0>(a, T, src, n0, stream);
    case 2 * 16 + 1: return launch_cd<2, 1, 0>(a, T, src, n0, stream);
    case 2 * 16 + 2: return launch_cd<2, 2, 0>(a, T, src, n0, stream);
    case 4 * 16 + 1: return launch_cd<4, 1, 0>(a, T, src, n0, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace fdtd2d
