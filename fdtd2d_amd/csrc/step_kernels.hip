// step_kernels.hip - fused, temporally blocked TMz step kernels for gfx950.
//
// One launch advances T (1..4) time steps. Mapping:
//  * a wavefront owns a tile: a strip of 64*C consecutive columns (C per
//    lane) and a band of H rows. Its outer HALO = roundup(T, C) columns on
//    each side are recomputed, not stored (VW = 64*C - 2*HALO are stored);
//  * the wave streams down its band one row per iteration. Level t (time step
//    n0+t) runs one row behind level t-1, so every level finds the rows it
//    needs in registers: Ez^{t-1} at rows r and r+1, Hx^{t-1/2} at r-1; the
//    x-neighbours (Ez at i+1, Hy at i-1) come from the adjacent lane;
//  * the band is extended by T rows above and below (the dependency cone of
//    T steps), clipped at the grid edge where the never-updated Ez ring
//    stops it;
//  * all state (Ez, Hx, Hy, Jz, the psi strips) is ping-ponged between two
//    buffer sets so that halo recomputation in neighbouring tiles reads the
//    step-n0 values while the owner writes step n0+T.
// Per cell and step the arithmetic is that of the reference kernels
// (F/kernel.cu:127-537, abbreviated as in fdtd2d.h), in the same order and
// without FMA contraction, so results match the CPU oracle bit for bit:
//   Hx = Hx + DB*((Ez[j]-Ez[j+1])*den_hy[j]);  + DB*psi_hxy     (y strips)
//   Hy = Hy + DB*((Ez[i+1]-Ez[i])*den_hx[i]);  + DB*psi_hyx     (x strips)
//   Jz = (float)(double(decay*Jz) + double((CJ*Ez^n)*Ne)*jz_dt)  (Ne > 0)
//   Ez = CA*Ez + CB*(((Hy[i]-Hy[i-1])*den_ex + (Hx[j-1]-Hx[j])*den_ey) - Jz)
//        (PEC material: 0), then + CB*psi_ezx, then + CB*psi_ezy
//   psi = b*psi + (c*diff)/d
// followed by the soft sources and the detector samples of that step.
#include <hip/hip_runtime.h>

#include "plan_internal.h"

#pragma clang fp contract(off)

namespace fdtd2d {
namespace {

constexpr int kTable = 0;
constexpr int kDensity = 1;

template <int C, bool PSI>
struct Level {
  float ez[C], hx[C], hy[C], jz[C];
  uint32_t med[C];  // TABLE: material id; DENSITY: bits of Ne
  float phxy[PSI ? C : 1], pezy[PSI ? C : 1], phyx[PSI ? C : 1], pezx[PSI ? C : 1];
};

__device__ __forceinline__ float jz_update(float jz, float decay, float cj, float ez, float ne,
                                           double kdt) {
  const float a = decay * jz;
  const float b = (cj * ez) * ne;
  return static_cast<float>(static_cast<double>(a) + static_cast<double>(b) * kdt);
}

// Column constants of one lane.
template <int C, bool PSI>
struct Cols {
  int c0;
  float den_ex[C], den_hx[C];
  float hyx_b[PSI ? C : 1], hyx_c[PSI ? C : 1], ezx_b[PSI ? C : 1], ezx_c[PSI ? C : 1];
  int hyx_s[PSI ? C : 1], ezx_s[PSI ? C : 1];
};

template <int C, bool PSI, int MED>
__device__ __forceinline__ void advance(const StepArgs &a, const float4 *lut_s, const Cols<C, PSI> &cl,
                                        const Level<C, PSI> &p, const float (&ez_s)[C],
                                        const float (&hx_n)[C], int r, int step, int own0, int own1,
                                        bool own_row, Level<C, PSI> &o) {
  const int nx = a.nx, ny = a.ny;
  const bool rowH = r >= 0 && r <= ny - 2;
  const bool rowE = r >= 1 && r <= ny - 2;
  const int rc = min(max(r, 0), ny - 2);
  const float den_hy = a.den_hy[rc];
  const float den_ey = a.den_ey[rc];
  int hxy_s = -1, ezy_s = -1;
  float hxy_b = 0.f, hxy_c = 0.f, ezy_b = 0.f, ezy_c = 0.f;
  if constexpr (PSI) {
    if (rowH) {
      hxy_s = a.hxy_slot[r];
      ezy_s = a.ezy_slot[r];
      hxy_b = a.hxy_b[r];
      hxy_c = a.hxy_c[r];
      ezy_b = a.ezy_b[r];
      ezy_c = a.ezy_c[r];
    }
  }
  // Ez^{t-1}(r, i+1)
  float ez_e[C];
#pragma unroll
  for (int c = 0; c < C - 1; ++c) ez_e[c] = p.ez[c + 1];
  ez_e[C - 1] = __shfl_down(p.ez[0], 1);
  // ---- H (F/kernel.cu:167-196, 198-240, 288-318, 321-362)
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int col = cl.c0 + c;
    const bool updH = rowH && col >= 0 && col <= nx - 2;
    float hx = p.hx[c], hy = p.hy[c];
    const float dEy = p.ez[c] - ez_s[c];
    const float dEx = ez_e[c] - p.ez[c];
    if (updH) {
      hx = hx + a.db * (dEy * den_hy);
      hy = hy + a.db * (dEx * cl.den_hx[c]);
    }
    if constexpr (PSI) {
      o.phxy[c] = p.phxy[c];
      o.phyx[c] = p.phyx[c];
      if (updH && hxy_s >= 0) {
        const float ps = hxy_b * p.phxy[c] + (hxy_c * dEy) / a.dyf;
        o.phxy[c] = ps;
        hx = hx + a.db * ps;
      }
      if (updH && cl.hyx_s[c] >= 0) {
        const float ps = cl.hyx_b[c] * p.phyx[c] + (cl.hyx_c[c] * dEx) / a.dxf;
        o.phyx[c] = ps;
        hy = hy + a.db * ps;
      }
    }
    o.hx[c] = hx;
    o.hy[c] = hy;
  }
  // Hy^{t-1/2}(r, i-1)
  float hy_w[C];
#pragma unroll
  for (int c = 1; c < C; ++c) hy_w[c] = o.hy[c - 1];
  hy_w[0] = __shfl_up(o.hy[C - 1], 1);
  // ---- Jz (F/kernel.cu:127-163) and Ez (F/kernel.cu:413-537)
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int col = cl.c0 + c;
    const bool updE = rowE && col >= 1 && col <= nx - 2;
    float ca, cb, ne;
    bool pec = false;
    if constexpr (MED == kTable) {
      const float4 m = lut_s[p.med[c]];
      ca = m.x;
      cb = m.y;
      ne = m.z;
      pec = static_cast<int>(p.med[c]) == a.pec;
    } else {
      ca = a.ca0;
      cb = a.cb0;
      ne = __uint_as_float(p.med[c]);
    }
    float jz = p.jz[c];
    if (updE && ne > 0.f) jz = jz_update(jz, a.jz_decay, a.cj, p.ez[c], ne, a.jz_dt);
    float ez = p.ez[c];
    const float dHy = o.hy[c] - hy_w[c];
    const float dHx = hx_n[c] - o.hx[c];
    if (updE) {
      if (pec)
        ez = 0.f;
      else
        ez = ca * ez + cb * ((dHy * cl.den_ex[c] + dHx * den_ey) - jz);
    }
    if constexpr (PSI) {
      o.pezx[c] = p.pezx[c];
      o.pezy[c] = p.pezy[c];
      if (updE && cl.ezx_s[c] >= 0) {
        const float ps = cl.ezx_b[c] * p.pezx[c] + (cl.ezx_c[c] * dHy) / a.dxf;
        o.pezx[c] = ps;
        ez = ez + cb * ps;
      }
      if (updE && ezy_s >= 0) {
        const float ps = ezy_b * p.pezy[c] + (ezy_c * dHx) / a.dyf;
        o.pezy[c] = ps;
        ez = ez + cb * ps;
      }
    }
    o.jz[c] = jz;
    o.ez[c] = ez;
    o.med[c] = p.med[c];
  }
  // ---- soft sources, then detector samples of this step (F/kernel.cu:621-805)
  if (r >= 0 && r < ny) {
    const int e0 = a.ev_off[r], e1 = a.ev_off[r + 1];
    for (int e = e0; e < e1; ++e) {
      const Event ev = a.events[e];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        if (cl.c0 + c != ev.i) continue;
        if (ev.kind == EV_SRC) {
          o.ez[c] = o.ez[c] - a.src_table[static_cast<long long>(step) * a.nsrc + ev.idx];
        } else if (own_row && ev.i >= own0 && ev.i < own1) {
          if (ev.kind == EV_DET) {
            a.det[static_cast<long long>(step) * a.ndet + ev.idx] = o.ez[c];
          } else {
            const long long q = static_cast<long long>(step) * a.ncomp + ev.idx;
            a.comp[q] = o.ez[c];
            a.comp[a.comp_plane + q] = o.hx[c];
            a.comp[2 * a.comp_plane + q] = o.hy[c];
          }
        }
      }
    }
  }
}

template <int C>
__device__ __forceinline__ void load_row(const float *__restrict__ base, int k, int c0, int pitch,
                                         bool ok, float (&v)[C]) {
  if (ok) {
    const float *q = base + static_cast<long long>(k) * pitch + c0;
    if constexpr (C == 1) {
      v[0] = q[0];
    } else if constexpr (C == 2) {
      const float2 t = *reinterpret_cast<const float2 *>(q);
      v[0] = t.x;
      v[1] = t.y;
    } else {
      const float4 t = *reinterpret_cast<const float4 *>(q);
      v[0] = t.x;
      v[1] = t.y;
      v[2] = t.z;
      v[3] = t.w;
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c) v[c] = 0.f;
  }
}

template <int T, int C, bool PSI, int MED>
__device__ __forceinline__ void run_tile(const StepArgs &a, const float4 *lut_s, int src, int n0,
                                         int tx, int ty) {
  constexpr int HALO = ((T + C - 1) / C) * C;
  constexpr int VW = 64 * C - 2 * HALO;
  const int lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny, pitch = a.pitch;
  const int cs = tx * VW - HALO;
  Cols<C, PSI> cl;
  cl.c0 = cs + lane * C;
  const int own0 = tx * VW, own1 = min(nx, own0 + VW);
  const int jb0 = ty * a.H, jb1 = min(ny, jb0 + a.H);
  const int jl0 = max(0, jb0 - T), jl1 = min(ny, jb1 + T);
  const bool lane_in = cl.c0 >= 0 && cl.c0 < pitch;  // the lane's C columns are in memory
  const float *ezA = a.ez[src], *hxA = a.hx[src], *hyA = a.hy[src], *jzA = a.jz[src];
  float *ezB = a.ez[1 - src], *hxB = a.hx[1 - src], *hyB = a.hy[1 - src], *jzB = a.jz[1 - src];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int col = cl.c0 + c;
    const bool in = col >= 0 && col < pitch;
    cl.den_ex[c] = in ? a.den_ex[col] : 0.f;
    cl.den_hx[c] = in ? a.den_hx[col] : 0.f;
    if constexpr (PSI) {
      cl.hyx_s[c] = in ? a.hyx_slot[col] : -1;
      cl.ezx_s[c] = in ? a.ezx_slot[col] : -1;
      cl.hyx_b[c] = in ? a.hyx_b[col] : 0.f;
      cl.hyx_c[c] = in ? a.hyx_c[col] : 0.f;
      cl.ezx_b[c] = in ? a.ezx_b[col] : 0.f;
      cl.ezx_c[c] = in ? a.ezx_c[col] : 0.f;
    }
  }
  Level<C, PSI> st[T + 1];
#pragma unroll
  for (int t = 0; t <= T; ++t)
#pragma unroll
    for (int c = 0; c < C; ++c) {
      st[t].ez[c] = st[t].hx[c] = st[t].hy[c] = st[t].jz[c] = 0.f;
      st[t].med[c] = 0u;
      if constexpr (PSI) st[t].phxy[c] = st[t].pezy[c] = st[t].phyx[c] = st[t].pezx[c] = 0.f;
    }

  for (int k = jl0; k < jl1 + T; ++k) {
    // ---- level 0: row k of the step-n0 state
    Level<C, PSI> in;
    const bool rv = k < jl1;
    const bool ld = rv && lane_in;
    load_row<C>(ezA, k, cl.c0, pitch, ld, in.ez);
    load_row<C>(hxA, k, cl.c0, pitch, ld, in.hx);
    load_row<C>(hyA, k, cl.c0, pitch, ld, in.hy);
    load_row<C>(jzA, k, cl.c0, pitch, ld, in.jz);
    {
      float mv[C];
      load_row<C>(MED == kTable ? reinterpret_cast<const float *>(a.med_id) : a.med_ne, k, cl.c0,
                  pitch, ld, mv);
#pragma unroll
      for (int c = 0; c < C; ++c) in.med[c] = __float_as_uint(mv[c]);
    }
    if constexpr (PSI) {
      const int sh = rv ? a.hxy_slot[k] : -1;
      const int se = rv ? a.ezy_slot[k] : -1;
      if (sh >= 0)
        load_row<C>(a.psi_hxy[src] + static_cast<long long>(sh) * pitch, 0, cl.c0, pitch, lane_in, in.phxy);
      else
#pragma unroll
        for (int c = 0; c < C; ++c) in.phxy[c] = 0.f;
      if (se >= 0)
        load_row<C>(a.psi_ezy[src] + static_cast<long long>(se) * pitch, 0, cl.c0, pitch, lane_in, in.pezy);
      else
#pragma unroll
        for (int c = 0; c < C; ++c) in.pezy[c] = 0.f;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const long long rowx = static_cast<long long>(k) * a.xs;
        in.phyx[c] = (rv && cl.hyx_s[c] >= 0) ? a.psi_hyx[src][rowx + cl.hyx_s[c]] : 0.f;
        in.pezx[c] = (rv && cl.ezx_s[c] >= 0) ? a.psi_ezx[src][rowx + cl.ezx_s[c]] : 0.f;
      }
    }
    Level<C, PSI> prev = st[0];
    st[0] = in;
    float nxt_ez[C];
#pragma unroll
    for (int c = 0; c < C; ++c) nxt_ez[c] = in.ez[c];
    // ---- levels 1..T, each one row behind the previous
#pragma unroll
    for (int t = 1; t <= T; ++t) {
      const int r = k - t;
      Level<C, PSI> out;
      const bool own_row = r >= jb0 && r < jb1;
      advance<C, PSI, MED>(a, lut_s, cl, prev, nxt_ez, st[t].hx, r, n0 + t, own0, own1, own_row, out);
      prev = st[t];
      st[t] = out;
#pragma unroll
      for (int c = 0; c < C; ++c) nxt_ez[c] = out.ez[c];
    }
    // ---- store step n0+T for owned cells of row k-T
    const int r = k - T;
    if (r >= jb0 && r < jb1) {
      const long long rowp = static_cast<long long>(r) * pitch;
      const Level<C, PSI> &o = st[T];
      int sh = -1, se = -1;
      if constexpr (PSI) {
        sh = a.hxy_slot[r];
        se = a.ezy_slot[r];
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int col = cl.c0 + c;
        if (col < own0 || col >= own1) continue;
        ezB[rowp + col] = o.ez[c];
        hxB[rowp + col] = o.hx[c];
        hyB[rowp + col] = o.hy[c];
        jzB[rowp + col] = o.jz[c];
        if constexpr (PSI) {
          if (sh >= 0) a.psi_hxy[1 - src][static_cast<long long>(sh) * pitch + col] = o.phxy[c];
          if (se >= 0) a.psi_ezy[1 - src][static_cast<long long>(se) * pitch + col] = o.pezy[c];
          const long long rowx = static_cast<long long>(r) * a.xs;
          if (cl.hyx_s[c] >= 0) a.psi_hyx[1 - src][rowx + cl.hyx_s[c]] = o.phyx[c];
          if (cl.ezx_s[c] >= 0) a.psi_ezx[1 - src][rowx + cl.ezx_s[c]] = o.pezx[c];
        }
      }
    }
  }
}

template <int T, int C, int MED>
__global__ void __launch_bounds__(64 * kWavesPerWG) step_kernel(const StepArgs a, int src, int n0) {
  __shared__ float4 lut_s[MED == kTable ? 256 : 1];
  if constexpr (MED == kTable) {
    for (int q = threadIdx.x; q < a.nmat; q += blockDim.x) lut_s[q] = a.lut[q];
    __syncthreads();
  }
  // Work-group -> (column group, band): consecutive work-groups of one XCD
  // (blockIdx % 8 round robin) take vertically adjacent bands so that the
  // band halos they share stay in that XCD's L2.
  const int G = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, q = G >> 3, rr = G & 7;
  const int L = xcd * q + min(xcd, rr) + (b >> 3);
  const int ty = L % a.nty;
  const int wgx = L / a.nty;
  const int tx = wgx * kWavesPerWG + (threadIdx.x >> 6);
  if (tx >= a.ntx) return;
  constexpr int HALO = ((T + C - 1) / C) * C;
  constexpr int VW = 64 * C - 2 * HALO;
  bool psi = false;
  if (a.P > 0) {
    const int cs = tx * VW - HALO, ce = cs + 64 * C;  // loaded columns [cs, ce)
    const int jb0 = ty * a.H, jb1 = min(a.ny, jb0 + a.H);
    const int jl0 = max(0, jb0 - T), jl1 = min(a.ny, jb1 + T);
    psi = cs <= a.P || ce >= a.nx - a.P - 1 || jl0 <= a.P || jl1 >= a.ny - a.P - 1;
  }
  if (psi)
    run_tile<T, C, true, MED>(a, lut_s, src, n0, tx, ty);
  else
    run_tile<T, C, false, MED>(a, lut_s, src, n0, tx, ty);
}

template <int T, int C>
hipError_t launch_tc(const StepArgs &a, int src, int n0, hipStream_t s) {
  const int grid = a.nwgx * a.nty;
  if (a.medium_kind == kTable)
    hipLaunchKernelGGL((step_kernel<T, C, kTable>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, a, src, n0);
  else
    hipLaunchKernelGGL((step_kernel<T, C, kDensity>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, a, src, n0);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_step(const StepArgs &a, int T, int C, int src, int n0, hipStream_t stream) {
  if (C != 1) return hipErrorInvalidValue;
  switch (T) {
    case 1: return launch_tc<1, 1>(a, src, n0, stream);
    case 2: return launch_tc<2, 1>(a, src, n0, stream);
    case 3: return launch_tc<3, 1>(a, src, n0, stream);
    case 4: return launch_tc<4, 1>(a, src, n0, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace fdtd2d
