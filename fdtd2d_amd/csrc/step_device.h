// step_device.h - fused, temporally blocked TMz step kernels for gfx950: the
// device code shared by the single-grid and pool kernel translation units
// (step_kernels.hip, fused_kernels.hip, pool_kernels.hip). HIP-only.
//
// One launch advances T (1..8) time steps. Mapping:
//  * a wavefront owns a tile: a strip of 64*C consecutive columns (C per
//    lane) and a band of rows. Its outer HALO >= T columns on each side are
//    recomputed, not stored (VW = 64*C - 2*HALO are stored);
//  * the wave streams down its band one row per iteration. Level t (time step
//    n0+t) runs one row behind level t-1, so every level finds the rows it
//    needs in registers: Ez^{t-1} at rows r and r+1, Hx^{t-1/2} at r-1; the
//    x-neighbours (Ez at i+1, Hy at i-1) come from the adjacent lane;
//  * the band is extended by T rows above and below (the dependency cone of
//    T steps), clipped at the grid edge where the never-updated Ez ring
//    stops it;
//  * all state (Ez, Hx, Hy, Jz, the psi strips) is ping-ponged between two
//    buffer sets so that halo recomputation in neighbouring tiles reads the
//    step-n0 values while the owner writes step n0+T;
//  * two kernels share a launch. CPML-class tiles (the tile loads a cell of a
//    CPML strip or of the outer ring, a row or column with non-uniform den,
//    or a source / detector / probe cell) carry what their window needs, in
//    three kinds compiled separately: x-strip tiles (per-column tables, the
//    x-strip psi and the reference's column ranges: the left and right
//    margins of the grid), y-strip tiles (per-row records, the y-strip psi
//    and the row ranges: the top and bottom bands) and general tiles (both,
//    plus the events: the corners and the windows around event cells).
//    Interior tiles carry none of these, so their per-cell code is
//    straight-line (no lane-divergent branches, no per-row scalars), and
//    they address HBM with one 32-bit row offset shared by all streams.
//    Interior tiles whose window holds no plasma and no Jz skip the Jz
//    update and its loads and stores.
// Per cell and step the arithmetic is that of the reference kernels
// (F/kernel.cu:127-537, abbreviated as in fdtd2d.h), in the same order and
// without FMA contraction, so results match the CPU oracle bit for bit:
//   Hx = Hx + DB*((Ez[j]-Ez[j+1])*den_hy[j]);  + DB*psi_hxy     (y strips)
//   Hy = Hy + DB*((Ez[i+1]-Ez[i])*den_hx[i]);  + DB*psi_hyx     (x strips)
//   Jz = (float)(double(decay*Jz) + double((CJ*Ez^n)*Ne)*jz_dt)  (Ne > 0)
//   Ez = CA*Ez + CB*(((Hy[i]-Hy[i-1])*den_ex + (Hx[j-1]-Hx[j])*den_ey) - Jz)
//        (PEC material: 0), then + CB*psi_ezx, then + CB*psi_ezy
//   psi = b*psi + (c*diff)/d   (float quotient, see div_by)
// followed by the soft sources and the detector samples of that step.
#ifndef FDTD2D_STEP_DEVICE_H
#define FDTD2D_STEP_DEVICE_H

#include <hip/hip_runtime.h>

#include "plan_internal.h"

#pragma clang fp contract(off)

namespace fdtd2d {
namespace {

constexpr int kTable = 0;
constexpr int kDensity = 1;

// What a tile carries (bits of K; interior tiles are K = 0):
constexpr int kX = 1;  // per-column den and CPML tables, x-strip psi, the reference's column ranges
constexpr int kY = 2;  // per-row records (den, y-strip psi slots), the reference's row ranges
constexpr int kE = 4;  // sources, detectors, component detectors, RC probes, processing line
constexpr int kGeneral = kX | kY | kE;

// Read-only tables (written by the host before any launch) are read through
// the constant address space, so uniform indices become scalar loads; plain
// global pointers would compile to vector loads plus s_waitcnt vmcnt(0),
// which also waits for the row prefetches in flight.
typedef const int __attribute__((address_space(4))) *cint_ptr;
typedef const float __attribute__((address_space(4))) *cfloat_ptr;
__device__ __forceinline__ int cld(const int *p, long long i) { return ((cint_ptr)p)[i]; }
__device__ __forceinline__ float cld(const float *p, long long i) { return ((cfloat_ptr)p)[i]; }
__device__ __forceinline__ int cld_i(const void *p, long long i) { return ((cint_ptr)p)[i]; }

// Device pointers held in StepArgs are plain pointers; dereferences go through
// the global address space explicitly, so that pointers read from a pool's
// descriptor table (not kernel arguments) still give global, not flat, memory
// instructions.
typedef float __attribute__((address_space(1))) gfloat;
typedef int __attribute__((address_space(1))) gint;
typedef char __attribute__((address_space(1))) gchar;
typedef float v2f __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef v2f __attribute__((address_space(1))) gv2f;
typedef v4f __attribute__((address_space(1))) gv4f;
__device__ __forceinline__ gfloat *gp(float *p) { return (gfloat *)p; }
__device__ __forceinline__ const gfloat *gp(const float *p) { return (const gfloat *)p; }
__device__ __forceinline__ const gint *gp(const int *p) { return (const gint *)p; }
__device__ __forceinline__ float4 gload4(const float4 *p, long long i) {
  const v4f t = ((const gv4f *)p)[i];
  return make_float4(t.x, t.y, t.z, t.w);
}

// x-neighbours across lanes with wavefront-wide DPP shifts (GFX9 wave_shl:1
// and wave_shr:1): lane l receives lane l+1 / l-1. The end lanes receive 0;
// they lie in the tile's halo, whose columns the levels give up one by one.
__device__ __forceinline__ float lane_from_next(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_from_prev(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}

template <int C, int K>
struct Level {
  float ez[C], hx[C], hy[C], jz[C];
  uint32_t med[C];  // TABLE: material id; DENSITY: bits of Ne
  float phxy[(K & kY) ? C : 1], pezy[(K & kY) ? C : 1], phyx[(K & kX) ? C : 1], pezx[(K & kX) ? C : 1];
};

// x / d for the plan's constant cell size d, correctly rounded as the float
// division of the reference: the product of x with the double nearest 1/d,
// rounded once to float. Unless the exact quotient is a float, it lies at
// least 2^-48 (relative) away from every midpoint between floats (x and d
// have 24-bit significands; when d is a power of two the double product is
// exact), and the double product is within 2^-52 of it, so both round to
// the same float. Checked against the float division for all 2^32 x at
// fourteen divisors, including subnormal ones.
__device__ __forceinline__ float div_by(float x, double rd) {
  return static_cast<float>(static_cast<double>(x) * rd);
}

__device__ __forceinline__ float jz_update(float jz, float decay, float cj, float ez, float ne,
                                           double kdt) {
  const float a = decay * jz;
  const float b = (cj * ez) * ne;
  return static_cast<float>(static_cast<double>(a) + static_cast<double>(b) * kdt);
}

// Column constants of one lane.
template <int C, int K>
struct Cols {
  int c0;
  static constexpr int NX = (K & kX) ? C : 1;
  float den_ex[NX], den_hx[NX];  // without kX: uniform scalars
  float hyx_b[NX], hyx_c[NX], ezx_b[NX], ezx_c[NX];
  int hyx_s[NX], ezx_s[NX];
  int probe[(K & kE) ? C : 1];  // RC capture probe of the column, -1 none
  bool wave_x;                  // some lane of the wave has an x-strip column
};

// Per-row scalars of the row a level works on (uniform across the wave):
// one 48-byte record per row, read by each level when it reaches the row
// (carrying them down the levels instead would hold T+1 records in SGPRs).
struct RowS {
  float den_hy, den_ey;
  int e0, e1;
  int hxy_s, ezy_s;
  float hxy_b, hxy_c, ezy_b, ezy_c;
};

typedef int rec_v4i __attribute__((ext_vector_type(4)));
typedef const rec_v4i __attribute__((address_space(4))) *rec_ptr;

// Record of row r; levels() passes rec_base = the record of row k - T, so
// that level t's record is a constant offset from it (one address per row
// iteration instead of one per level).
__device__ __forceinline__ rec_ptr row_record(const StepArgs &a, int r) {
  return (rec_ptr)a.rowinfo + (kRowRec / 4) * r;
}

template <int K>
__device__ __forceinline__ RowS row_scalars(const StepArgs &a, rec_ptr rec) {
  RowS q;
  q.hxy_s = q.ezy_s = -1;
  q.hxy_b = q.hxy_c = q.ezy_b = q.ezy_c = 0.f;
  if constexpr (!(K & (kY | kE))) {
    // tiles without per-row records: every row of their window has the
    // uniform den, no strip and no event
    q.den_hy = a.uden_hy;
    q.den_ey = a.uden_ey;
    q.e0 = q.e1 = 0;
    return q;
  }
  // (a tile with events but no kY reads the record's first 16 bytes for its
  // event range; its rows hold the uniform den and no strip)
  // the whole record in three 16-byte scalar loads issued together (loads of
  // single fields would be sunk into the branches that use them, each then
  // waiting for its own load); the records cover rows -kMaxT .. ny+kMaxT-1,
  // those outside the grid with no events and no strips
  const rec_v4i w0 = rec[0];
  q.den_hy = __int_as_float(w0.x);
  q.den_ey = __int_as_float(w0.y);
  q.e0 = w0.z >= 0 ? w0.z : 0;
  q.e1 = w0.z >= 0 ? w0.w : 0;
  if constexpr (!(K & kY)) return q;  // event tiles: den and the event range only
  const rec_v4i w1 = rec[1], w2 = rec[2];
  q.hxy_s = w1.x;
  q.ezy_s = w1.y;
  q.hxy_b = __int_as_float(w1.z);
  q.hxy_c = __int_as_float(w1.w);
  q.ezy_b = __int_as_float(w2.x);
  q.ezy_c = __int_as_float(w2.y);
  return q;
}

// One level: level t-1's row r (p), Ez^{t-1}(r+1) (ez_s) and Hx^{t-1/2}(r-1)
// (hx_n) -> level t's row r (o).
template <int C, int K, int MED, int JZ, bool XT>
__device__ __forceinline__ void advance(const StepArgs &a, const float4 *lut_s, const Cols<C, K> &cl,
                                        const RowS &rs, const Level<C, K> &p, const float (&ez_s)[C],
                                        const float (&hx_n)[C], int r, int step, int own0, int own1,
                                        bool own_row, Level<C, K> &o) {
  constexpr bool KX = (K & kX) != 0, KY = (K & kY) != 0, KE = (K & kE) != 0;
  const int nx = a.nx, ny = a.ny;
  // Index ranges of the reference kernels. A tile without kY loads only rows
  // inside the row ranges, one without kX only columns inside the column
  // ranges (the host checks each window), so there the masks are
  // compile-time true.
  const bool rowH = !KY || (r >= 0 && r <= ny - 2);
  const bool rowE = !KY || (r >= 1 && r <= ny - 2);
  // Ez^{t-1}(r, i+1)
  float ez_e[C];
#pragma unroll
  for (int c = 0; c < C - 1; ++c) ez_e[c] = p.ez[c + 1];
  ez_e[C - 1] = lane_from_next(p.ez[0]);
  // ---- H (F/kernel.cu:167-196, 198-240, 288-318, 321-362)
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int col = cl.c0 + c;
    const bool updH = rowH && (!KX || (col >= 0 && col <= nx - 2));
    const float dEy = p.ez[c] - ez_s[c];
    const float dEx = ez_e[c] - p.ez[c];
    float hx = p.hx[c] + a.db * (dEy * rs.den_hy);
    float den_hx;
    if constexpr (KX) den_hx = cl.den_hx[c];
    else den_hx = a.uden_hx;
    float hy = p.hy[c] + a.db * (dEx * den_hx);
    if constexpr (K != 0) {
      hx = updH ? hx : p.hx[c];
      hy = updH ? hy : p.hy[c];
    }
    // psi strips: the branches test wave-uniform conditions (the row's strip
    // slot, or some lane of the wave in an x strip) and the lanes select, so
    // no branch splits the wave
    if constexpr (KY) {
      o.phxy[c] = p.phxy[c];
      if (rs.hxy_s >= 0) {
        const float ps = rs.hxy_b * p.phxy[c] + div_by(rs.hxy_c * dEy, a.rdy);
        o.phxy[c] = updH ? ps : p.phxy[c];
        hx = updH ? hx + a.db * ps : hx;
      }
    }
    if constexpr (KX) {
      o.phyx[c] = p.phyx[c];
      if (cl.wave_x) {
        const bool u = updH && cl.hyx_s[c] >= 0;
        const float ps = cl.hyx_b[c] * p.phyx[c] + div_by(cl.hyx_c[c] * dEx, a.rdx);
        o.phyx[c] = u ? ps : p.phyx[c];
        hy = u ? hy + a.db * ps : hy;
      }
    }
    o.hx[c] = hx;
    o.hy[c] = hy;
  }
  // Hy^{t-1/2}(r, i-1)
  float hy_w[C];
#pragma unroll
  for (int c = 1; c < C; ++c) hy_w[c] = o.hy[c - 1];
  hy_w[0] = lane_from_prev(o.hy[C - 1]);
  // ---- Jz (F/kernel.cu:127-163) and Ez (F/kernel.cu:413-537)
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int col = cl.c0 + c;
    const bool updE = rowE && (!KX || (col >= 1 && col <= nx - 2));
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
    const float dHy = o.hy[c] - hy_w[c];
    const float dHx = hx_n[c] - o.hx[c];
    float den_ex;
    if constexpr (KX) den_ex = cl.den_ex[c];
    else den_ex = a.uden_ex;
    const float curl = dHy * den_ex + dHx * rs.den_ey;
    float jz = 0.f, ez;
    if constexpr (JZ != 0) {
      // Jz^{n+1/2} from Ez^n where Ne > 0; (...) - Jz as the reference orders it
      const float jz_new = jz_update(p.jz[c], a.jz_decay, a.cj, p.ez[c], ne, a.jz_dt);
      // JZ == 2: every cell of the tile's window has Ne > 0 (checked on the
      // host), so the test always picks the update
      jz = (JZ == 2 || (updE && ne > 0.f)) ? jz_new : p.jz[c];
      ez = ca * p.ez[c] + cb * (curl - jz);
    } else {
      // the window holds no plasma and Jz = +0, so curl - Jz = curl exactly
      (void)ne;
      ez = ca * p.ez[c] + cb * curl;
    }
    if constexpr (MED == kTable) ez = pec ? 0.f : ez;
    if constexpr (K != 0) ez = updE ? ez : p.ez[c];
    // +psi_x, then +psi_y (corners get both, in the reference's order)
    if constexpr (KX) {
      o.pezx[c] = p.pezx[c];
      if (cl.wave_x) {
        const bool u = updE && cl.ezx_s[c] >= 0;
        const float ps = cl.ezx_b[c] * p.pezx[c] + div_by(cl.ezx_c[c] * dHy, a.rdx);
        o.pezx[c] = u ? ps : p.pezx[c];
        ez = u ? ez + cb * ps : ez;
      }
    }
    if constexpr (KY) {
      o.pezy[c] = p.pezy[c];
      if (rs.ezy_s >= 0) {
        const float ps = rs.ezy_b * p.pezy[c] + div_by(rs.ezy_c * dHx, a.rdy);
        o.pezy[c] = updE ? ps : p.pezy[c];
        ez = updE ? ez + cb * ps : ez;
      }
    }
    o.jz[c] = jz;
    o.ez[c] = ez;
    o.med[c] = p.med[c];
  }
  if constexpr (!KE) return;
  // ---- soft sources, then detector samples of this step (F/kernel.cu:621-805)
  for (int e = rs.e0; e < rs.e1; ++e) {
    Event ev;
    ev.i = cld_i(a.events, 4LL * e);
    ev.kind = cld_i(a.events, 4LL * e + 1);
    ev.idx = cld_i(a.events, 4LL * e + 2);
#pragma unroll
    for (int c = 0; c < C; ++c) {
      if (cl.c0 + c != ev.i) continue;
      if (ev.kind == EV_SRC) {
        o.ez[c] = o.ez[c] - cld(a.src_table, static_cast<long long>(step) * a.nsrc + ev.idx);
      } else if (own_row && ev.i >= own0 && ev.i < own1) {
        if (ev.kind == EV_DET) {
          gp(a.det)[static_cast<long long>(step) * a.ndet + ev.idx] = o.ez[c];
        } else {
          const long long q = static_cast<long long>(step) * a.ncomp + ev.idx;
          gp(a.comp)[q] = o.ez[c];
          gp(a.comp)[a.comp_plane + q] = o.hx[c];
          gp(a.comp)[2 * a.comp_plane + q] = o.hy[c];
        }
      }
    }
  }
  if constexpr (XT) {
    // RC capture of probe columns (after the sources, as the reference's
    // Dev_RCcalc reads Ez after the detectors of the step)
    if (a.nprobe > 0 && own_row) {
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int col = cl.c0 + c;
        if (cl.probe[c] >= 0 && col >= own0 && col < own1)
          gp(a.rc_cap)[(static_cast<long long>((step - 1) & (kMaxT - 1)) * a.nprobe + cl.probe[c]) * a.ny + r] =
              o.ez[c];
      }
    }
    // processing line (Dev_SaveLine runs after the sources, before the
    // detectors; all three read the same Ez)
    if (r == a.line_j && own_row && step > a.line_begin && step <= a.line_end) {
      const long long t = step - a.line_begin - 1;
      const int w = a.line_i1 - a.line_i0;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int col = cl.c0 + c;
        if (col >= a.line_i0 && col < a.line_i1 && col >= own0 && col < own1)
          gp(a.line)[t * w + (col - a.line_i0)] = o.ez[c];
      }
    }
  }
}

// ---------------------------------------------------------------- memory
// Interior tiles address every stream as (uniform base pointer) + (32-bit
// byte offset shared by all streams): one VGPR offset per row, 2 SGPRs per
// stream, and the loads take the global_load v, voff, s[base] form.
struct Streams {
  const gchar *ez, *hx, *hy, *jz, *med;
};

template <int C>
__device__ __forceinline__ void oload(const gchar *base, unsigned off, float (&v)[C]) {
  const gchar *q = base + off;
  if constexpr (C == 1) {
    v[0] = *(const gfloat *)q;
  } else if constexpr (C == 2) {
    const v2f t = *(const gv2f *)q;
    v[0] = t.x;
    v[1] = t.y;
  } else {
    const v4f t = *(const gv4f *)q;
    v[0] = t.x;
    v[1] = t.y;
    v[2] = t.z;
    v[3] = t.w;
  }
}

template <int C>
__device__ __forceinline__ void ostore(const gchar *base, unsigned off, const float (&v)[C]) {
  gchar *q = const_cast<gchar *>(base) + off;
  if constexpr (C == 1) {
    *(gfloat *)q = v[0];
  } else if constexpr (C == 2) {
    *(gv2f *)q = v2f{v[0], v[1]};
  } else {
    *(gv4f *)q = v4f{v[0], v[1], v[2], v[3]};
  }
}

template <int C>
__device__ __forceinline__ void load_row(const float *base_, int k, int c0, int pitch,
                                         bool ok, float (&v)[C]) {
  if (ok) {
    const gfloat *q = gp(base_) + static_cast<long long>(k) * pitch + c0;
    if constexpr (C == 1) {
      v[0] = q[0];
    } else if constexpr (C == 2) {
      const v2f t = *(const gv2f *)q;
      v[0] = t.x;
      v[1] = t.y;
    } else {
      const v4f t = *(const gv4f *)q;
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

// Stores the C values of a lane at columns [c0, c0+C) of row r, only those in
// [own0, own1): one vector store when all are owned.
template <int C>
__device__ __forceinline__ void store_row(float *base_, long long rowp, int c0, int own0, int own1,
                                          const float (&v)[C]) {
  gfloat *q = gp(base_) + rowp + c0;
  if (c0 >= own0 && c0 + C <= own1) {
    if constexpr (C == 1) {
      q[0] = v[0];
    } else if constexpr (C == 2) {
      *(gv2f *)q = v2f{v[0], v[1]};
    } else {
      *(gv4f *)q = v4f{v[0], v[1], v[2], v[3]};
    }
  } else {
#pragma unroll
    for (int c = 0; c < C; ++c)
      if (c0 + c >= own0 && c0 + c < own1) q[c] = v[c];
  }
}

// Level 0 of row k: the step-n0 state (zeros for rows outside [0, ny) or
// past the band's last loaded row).
template <int C, int K, int MED, int JZ>
__device__ __forceinline__ void load_level0(const StepArgs &a, const Streams &sm, const Cols<C, K> &cl,
                                            int src, int k, bool rv, bool lane_in, Level<C, K> &in) {
  const int pitch = a.pitch;
  if constexpr (K == 0) {
    // interior: every lane's columns are inside the grid. The loads are
    // unconditional (the caller clamps k to the band's last loaded row; rows
    // past it lie outside the dependency cone of the stored rows), so every
    // iteration issues the same memory operations and the wait before the
    // next row's use covers exactly that row's loads.
    (void)rv;
    const unsigned off = static_cast<unsigned>(cl.c0) * 4u + static_cast<unsigned>(k) * static_cast<unsigned>(pitch) * 4u;
    float mv[C];
    oload<C>(sm.ez, off, in.ez);
    oload<C>(sm.hx, off, in.hx);
    oload<C>(sm.hy, off, in.hy);
    if constexpr (JZ != 0) {
      oload<C>(sm.jz, off, in.jz);
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) in.jz[c] = 0.f;
    }
    if constexpr (JZ != 0 || MED == kTable) {
      oload<C>(sm.med, off, mv);
#pragma unroll
      for (int c = 0; c < C; ++c) in.med[c] = __float_as_uint(mv[c]);
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) in.med[c] = 0u;
    }
  } else {
    const bool ld = rv && lane_in;
    load_row<C>(a.ez[src], k, cl.c0, pitch, ld, in.ez);
    load_row<C>(a.hx[src], k, cl.c0, pitch, ld, in.hx);
    load_row<C>(a.hy[src], k, cl.c0, pitch, ld, in.hy);
    load_row<C>(a.jz[src], k, cl.c0, pitch, ld, in.jz);
    {
      float mv[C];
      load_row<C>(MED == kTable ? reinterpret_cast<const float *>(a.med_id) : a.med_ne, k, cl.c0, pitch, ld, mv);
#pragma unroll
      for (int c = 0; c < C; ++c) in.med[c] = __float_as_uint(mv[c]);
    }
    if constexpr ((K & kY) != 0) {
      const int sh = rv ? cld(a.hxy_slot, k) : -1;
      const int se = rv ? cld(a.ezy_slot, k) : -1;
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
    }
    if constexpr ((K & kX) != 0) {
      const long long rowx = static_cast<long long>(k) * a.xs;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        in.phyx[c] = (rv && cl.hyx_s[c] >= 0) ? gp(a.psi_hyx[src])[rowx + cl.hyx_s[c]] : 0.f;
        in.pezx[c] = (rv && cl.ezx_s[c] >= 0) ? gp(a.psi_ezx[src])[rowx + cl.ezx_s[c]] : 0.f;
      }
    }
  }
}

// Levels 1..T of one row iteration: level t computes row k-t from level t-1.
// General tiles (kE) apply the sources and take the samples of the events on
// those rows (only they hold event cells).
// WARM (the first iterations of a band): level t's rows are needed from row
// jb0 - (T-t) - 1 on (the rows the stored rows depend on, plus one row of Hx
// above them), i.e. from iteration k = jb0 - T - 1 + 2t; before that the
// level and all later ones are skipped. A skipped level leaves its pipeline
// register as it was; the first needed row of a level reads only its Hx
// from that register's rows, and only for an Ez nothing needs.
template <int T, int C, int K, int MED, int JZ, bool XT, bool WARM>
__device__ __forceinline__ void levels(const StepArgs &a, const float4 *lut_s, const Cols<C, K> &cl, int n0,
                                       int k, int jb0, int jb1, int own0, int own1, Level<C, K> prev,
                                       Level<C, K> (&st)[T + 1]) {
  float nxt_ez[C];
#pragma unroll
  for (int c = 0; c < C; ++c) nxt_ez[c] = st[0].ez[c];
  const rec_ptr rec_base = row_record(a, k - T);  // k - T >= -kMaxT: the records' lower padding
#pragma unroll
  for (int t = 1; t <= T; ++t) {
    if constexpr (WARM) {
      if (k < jb0 - T - 1 + 2 * t) break;
    }
    const int r = k - t;
    Level<C, K> out;
    const bool own_row = r >= jb0 && r < jb1;
    const RowS rs = row_scalars<K>(a, rec_base + (kRowRec / 4) * (T - t));
    advance<C, K, MED, JZ, XT>(a, lut_s, cl, rs, prev, nxt_ez, st[t].hx, r, n0 + t, own0, own1, own_row, out);
    prev = st[t];
    st[t] = out;
#pragma unroll
    for (int c = 0; c < C; ++c) nxt_ez[c] = out.ez[c];
  }
}

// One row iteration: level 0 takes row k, level t computes row k-t, and
// row k-T is stored.
template <int T, int C, int K, int MED, int JZ, bool XT, bool WARM>
__device__ __forceinline__ void row_step(const StepArgs &a, const Streams &dst, const float4 *lut_s,
                                         const Cols<C, K> &cl, int src, int n0, int k, int jb0, int jb1,
                                         int own0, int own1, bool lane_owns, unsigned trash_off,
                                         const Level<C, K> &in, Level<C, K> (&st)[T + 1]) {
  Level<C, K> prev = st[0];
  st[0] = in;
  // ---- levels 1..T, each one row behind the previous
  levels<T, C, K, MED, JZ, XT, WARM>(a, lut_s, cl, n0, k, jb0, jb1, own0, own1, prev, st);
  // ---- store step n0+T for owned cells of row k-T
  const int r = k - T;
  if constexpr (K == 0) {
    // Unconditional stores (see load_level0): lanes or rows this tile does
    // not own write this wave's scratch slot below the grid instead.
    const bool ok = lane_owns && r >= jb0 && r < jb1;
    const unsigned off = ok ? static_cast<unsigned>(cl.c0) * 4u +
                                  static_cast<unsigned>(r) * static_cast<unsigned>(a.pitch) * 4u
                            : trash_off;
    const Level<C, K> &o = st[T];
    ostore<C>(dst.ez, off, o.ez);
    ostore<C>(dst.hx, off, o.hx);
    ostore<C>(dst.hy, off, o.hy);
    if constexpr (JZ != 0) ostore<C>(dst.jz, off, o.jz);
  } else if (r >= jb0 && r < jb1) {
    const int pitch = a.pitch;
    const Level<C, K> &o = st[T];
    const long long rowp = static_cast<long long>(r) * pitch;
    // the destination planes were resolved once per tile (dst), so the
    // row loop issues no scalar loads for them
    store_row<C>((float *)dst.ez, rowp, cl.c0, own0, own1, o.ez);
    store_row<C>((float *)dst.hx, rowp, cl.c0, own0, own1, o.hx);
    store_row<C>((float *)dst.hy, rowp, cl.c0, own0, own1, o.hy);
    store_row<C>((float *)dst.jz, rowp, cl.c0, own0, own1, o.jz);
    if constexpr ((K & kY) != 0) {
      const int sh = cld(a.hxy_slot, r), se = cld(a.ezy_slot, r);
      if (sh >= 0) store_row<C>(a.psi_hxy[1 - src], static_cast<long long>(sh) * pitch, cl.c0, own0, own1, o.phxy);
      if (se >= 0) store_row<C>(a.psi_ezy[1 - src], static_cast<long long>(se) * pitch, cl.c0, own0, own1, o.pezy);
    }
    if constexpr ((K & kX) != 0) {
      const long long rowx = static_cast<long long>(r) * a.xs;
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int col = cl.c0 + c;
        if (col < own0 || col >= own1) continue;
        if (cl.hyx_s[c] >= 0) gp(a.psi_hyx[1 - src])[rowx + cl.hyx_s[c]] = o.phyx[c];
        if (cl.ezx_s[c] >= 0) gp(a.psi_ezx[1 - src])[rowx + cl.ezx_s[c]] = o.pezx[c];
      }
    }
  }
}

// CPML-class tiles compile their row loop twice: with the RC probe capture
// and the processing line capture (XT), and without, for the launches of a
// problem that has neither (their per-level tests are about a tenth of the
// CPML-class kernel's instructions).
__device__ __forceinline__ bool extras_on(const StepArgs &a) {
  return a.nprobe > 0 || (a.line_i1 > a.line_i0 && a.line_end > a.line_begin);
}

// D = rows fetched ahead of the one being consumed (memory-level
// parallelism per wave); the row loop is unrolled by D so the prefetch ring
// is indexed statically.
// x: the interior entry's tile index (columns [x*vw, x*vw+vw)), or the
// CPML-class entry's packed first column and width (one column per lane).
// A CPML-class wave loads the 64 columns from halo_e (>= T) left of its
// first stored column, or from column 0: at the grid's left edge no halo is
// needed (column 0 depends on no column left of it), so the piece there
// stores up to 64 - T columns; at the right edge the host starts the piece
// 64 - T columns before nx, for the same reason.
template <int T, int C, int D, int K, int MED, int JZ, bool XT = false>
__device__ __forceinline__ void run_tile(const StepArgs &a, const float4 *lut_s, int src, int n0,
                                         int x, int jb0, int jb1) {
  static_assert(K == 0 || C == 1, "CPML-class tiles run one column per lane");
  constexpr bool KX = (K & kX) != 0, KE = (K & kE) != 0;
  const int lane = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny, pitch = a.pitch;
  int own0, own1, cs;
  if constexpr (K != 0) {
    own0 = x & ((1 << kEdgeColBits) - 1);
    own1 = min(nx, own0 + (x >> kEdgeColBits));
    cs = max(0, own0 - a.halo_e);  // halo_e >= T
  } else {
    own0 = a.x0 + x * a.vw;  // vw = 64*C - 2*halo, halo >= T and a multiple of C
    own1 = min(nx, own0 + a.vw);
    cs = own0 - a.halo;
  }
  Cols<C, K> cl;
  cl.c0 = cs + lane * C;
  const int jl0 = max(0, jb0 - T), jl1 = min(ny, jb1 + T);
  const bool lane_in = cl.c0 >= 0 && cl.c0 < pitch;  // the lane's C columns are in memory
  const bool lane_owns = cl.c0 >= own0 && cl.c0 + C <= own1;
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const int col = cl.c0 + c;
    const bool in = col >= 0 && col < pitch;
    if constexpr (KX) {
      cl.den_ex[c] = in ? gp(a.den_ex)[col] : 0.f;
      cl.den_hx[c] = in ? gp(a.den_hx)[col] : 0.f;
      cl.hyx_s[c] = in ? gp(a.hyx_slot)[col] : -1;
      cl.ezx_s[c] = in ? gp(a.ezx_slot)[col] : -1;
      cl.hyx_b[c] = in ? gp(a.hyx_b)[col] : 0.f;
      cl.hyx_c[c] = in ? gp(a.hyx_c)[col] : 0.f;
      cl.ezx_b[c] = in ? gp(a.ezx_b)[col] : 0.f;
      cl.ezx_c[c] = in ? gp(a.ezx_c)[col] : 0.f;
    }
    if constexpr (KE) cl.probe[c] = in ? gp(a.probe_col)[col] : -1;
  }
  cl.wave_x = false;
  if constexpr (KX) {
    bool x = false;
#pragma unroll
    for (int c = 0; c < C; ++c) x = x || cl.hyx_s[c] >= 0 || cl.ezx_s[c] >= 0;
    cl.wave_x = __builtin_amdgcn_ballot_w64(x) != 0;
  }
  Streams sm, dst;
  sm.ez = (const gchar *)a.ez[src];
  sm.hx = (const gchar *)a.hx[src];
  sm.hy = (const gchar *)a.hy[src];
  sm.jz = (const gchar *)a.jz[src];
  sm.med = MED == kTable ? (const gchar *)a.med_id : (const gchar *)a.med_ne;
  dst.ez = (const gchar *)a.ez[1 - src];
  dst.hx = (const gchar *)a.hx[1 - src];
  dst.hy = (const gchar *)a.hy[1 - src];
  dst.jz = (const gchar *)a.jz[1 - src];
  dst.med = sm.med;
  Level<C, K> st[T + 1];
#pragma unroll
  for (int t = 0; t <= T; ++t) {
#pragma unroll
    for (int c = 0; c < C; ++c) {
      st[t].ez[c] = st[t].hx[c] = st[t].hy[c] = st[t].jz[c] = 0.f;
      st[t].med[c] = 0u;
      if constexpr ((K & kY) != 0) st[t].phxy[c] = st[t].pezy[c] = 0.f;
      if constexpr (KX) st[t].phyx[c] = st[t].pezx[c] = 0.f;
    }
  }
  // Scratch slot of this wave in the rows below the grid (see row_step).
  const unsigned nslot = static_cast<unsigned>(a.trash_slots);
  const unsigned slot = (blockIdx.x * kWavesPerWG + (threadIdx.x >> 6)) % nslot;
  const unsigned trash_off =
      (static_cast<unsigned>(ny) * static_cast<unsigned>(pitch) + slot * 64u * C + static_cast<unsigned>(lane * C)) * 4u;
  // Drain the loop-invariant loads above: otherwise the row loop's wait
  // counts cannot separate them from the prefetches and wait for everything.
  __builtin_amdgcn_s_waitcnt(0);
  Level<C, K> pre[D];
#pragma unroll
  for (int u = 0; u < D; ++u) {
    const int kk = K != 0 ? jl0 + u : min(jl0 + u, jl1 - 1);
    load_level0<C, K, MED, JZ>(a, sm, cl, src, kk, jl0 + u < jl1, lane_in, pre[u]);
  }
  // Iterations k in [jl0, jb1 + T): the band's last row is stored at
  // k = jb1 - 1 + T. The first ones, up to k = jb0 + T - 2, run only the
  // levels whose rows are needed (WARM); the count is rounded up to a
  // multiple of D, and rows past the end store nothing. With more than one
  // column per lane a separate warm-up loop would double the live state the
  // compiler allocates for (occupancy 2); there one loop checks every level.
  const int kwarm = C > 1 ? jl0 : jb0 + T - 1, kend = jb1 + T;
  int k = jl0;
  for (; k < kwarm; k += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const Level<C, K> in = pre[u];
      const int kn = k + u + D;
      const int kk = K != 0 ? kn : min(kn, jl1 - 1);
      load_level0<C, K, MED, JZ>(a, sm, cl, src, kk, kn < jl1, lane_in, pre[u]);
      row_step<T, C, K, MED, JZ, XT, true>(a, dst, lut_s, cl, src, n0, k + u, jb0, jb1, own0, own1, lane_owns,
                                          trash_off, in, st);
    }
  }
  for (; k < kend; k += D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const Level<C, K> in = pre[u];
      const int kn = k + u + D;
      const int kk = K != 0 ? kn : min(kn, jl1 - 1);
      load_level0<C, K, MED, JZ>(a, sm, cl, src, kk, kn < jl1, lane_in, pre[u]);
      row_step<T, C, K, MED, JZ, XT, (C > 1)>(a, dst, lut_s, cl, src, n0, k + u, jb0, jb1, own0, own1, lane_owns,
                                             trash_off, in, st);
    }
  }
}

// A CPML-class entry by its kind (set by the host from the entry's window,
// see kKindGeneral .. kKindE). SET picks the kinds compiled: 0 every kind
// (the CPML-class kernel of an unfused launch); 1 the fusable kinds
// (fusable_kind), for the fused kernels: every kind for the density medium,
// for the table medium the strips and the event entries of problems without
// captures (with every kind that kernel would need 168 VGPRs and spill 96
// bytes per lane); 2 the rest, for the kernel beside a fused one. The row
// loops of general and event entries are compiled with the RC probe and
// processing-line captures and without, the latter for problems that have
// neither.
constexpr int cpml_set(int set, int med) { return set == 1 && fused_kinds(med) == kAllKinds ? 0 : set; }
constexpr bool set_has(int set, int med, int kind) {
  return set == 0 ? true
         : set == 1 ? ((fused_kinds(med) >> kind) & 1) != 0
                    : kind == kKindGeneral || kind == kKindE || !((fused_kinds(med) >> kind) & 1);
}

template <int T, int MED, int SET_ = 0>
__device__ __forceinline__ void run_cpml_tile(const StepArgs &a, const float4 *lut_s, int src, int n0, int x,
                                              int r0, int r1, int flags) {
  constexpr int SET = cpml_set(SET_, MED);
  switch (entry_kind(flags)) {
    case kKindX:
      if constexpr (set_has(SET, MED, kKindX)) run_tile<T, 1, 1, kX, MED, 1>(a, lut_s, src, n0, x, r0, r1);
      return;
    case kKindY:
      if constexpr (set_has(SET, MED, kKindY)) run_tile<T, 1, 1, kY, MED, 1>(a, lut_s, src, n0, x, r0, r1);
      return;
    case kKindXY:
      if constexpr (set_has(SET, MED, kKindXY)) run_tile<T, 1, 1, kX | kY, MED, 1>(a, lut_s, src, n0, x, r0, r1);
      return;
    case kKindE:
      if constexpr (SET == 1) {
        run_tile<T, 1, 1, kE, MED, 1, false>(a, lut_s, src, n0, x, r0, r1);
      } else if constexpr (SET == 2) {
        run_tile<T, 1, 1, kE, MED, 1, true>(a, lut_s, src, n0, x, r0, r1);
      } else {
        if (extras_on(a))
          run_tile<T, 1, 1, kE, MED, 1, true>(a, lut_s, src, n0, x, r0, r1);
        else
          run_tile<T, 1, 1, kE, MED, 1, false>(a, lut_s, src, n0, x, r0, r1);
      }
      return;
    default:
      if constexpr (set_has(SET, MED, kKindGeneral)) {
        if (extras_on(a))
          run_tile<T, 1, 1, kGeneral, MED, 1, true>(a, lut_s, src, n0, x, r0, r1);
        else
          run_tile<T, 1, 1, kGeneral, MED, 1, false>(a, lut_s, src, n0, x, r0, r1);
      }
      return;
  }
}

// Single-grid step kernel. MODE 0: the interior list (tile_list[0]); 1: the
// CPML-class list (tile_list[1]), every kind; 2: one list (tile_list[0])
// holding the interior entries and the CPML-class entries of the fusable
// kinds, each XCD's range starting with its share of the latter (flagged
// kFlagCpml), so that one launch covers the grid; 3: the other CPML-class
// entries (tile_list[1]: events inside a strip's window, say, or any event
// tile of a problem with captures; none in the 8192^2 workload), beside a
// MODE 2 kernel. The fused form is held to the 3 waves per SIMD the interior
// kernel runs at.
// The arguments are read through a constant-address-space pointer to the
// kernel argument segment (scalar loads): bound by value, the larger kernels
// copied the whole StepArgs to scratch and read it with vector loads.
typedef const StepArgs __attribute__((address_space(4))) *cargs_ptr;

template <int T, int C, int D, int MED, int MODE>
__global__ void __launch_bounds__(64 * kWavesPerWG, MODE == 2 ? 3 : 1)
    step_kernel(const StepArgs args, int src, int n0) {
  const StepArgs &a = *(const StepArgs *)((cargs_ptr)__builtin_amdgcn_kernarg_segment_ptr());
  __shared__ float4 lut_s[MED == kTable ? 256 : 1];
  if constexpr (MED == kTable) {
    for (int q = threadIdx.x; q < a.nmat; q += blockDim.x) lut_s[q] = gload4(a.lut, q);
    __syncthreads();
  }
  // Launch index -> list position: consecutive positions go to one XCD
  // (work-groups are dealt to the 8 XCDs round robin), and the lists are
  // ordered by column group then first row, so an XCD works down vertically
  // adjacent bands whose shared halo rows stay in its L2.
  const int G = gridDim.x, b = blockIdx.x;
  const int xcd = b & 7, q = G >> 3, rr = G & 7;
  const int pos = xcd * q + min(xcd, rr) + (b >> 3);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long e = 4LL * (pos * kWavesPerWG + wave);
  const int4 *list = a.tile_list[(MODE == 1 || MODE == 3) ? 1 : 0];
  const int tx = cld_i(list, e), r0 = cld_i(list, e + 1), r1 = cld_i(list, e + 2), flags = cld_i(list, e + 3);
  if (r0 >= r1) return;  // padding entry
  if constexpr (MODE == 1 || MODE == 3) {
    run_cpml_tile<T, MED, MODE == 1 ? 0 : 2>(a, lut_s, src, n0, tx, r0, r1, flags);
    return;
  } else {
    if constexpr (MODE == 2) {
      if (flags & kFlagCpml) {
        run_cpml_tile<T, MED, 1>(a, lut_s, src, n0, tx, r0, r1, flags);
        return;
      }
    }
    if (flags & 1) {
      // tiles full of plasma skip the Ne > 0 test (two columns per lane and a
      // density medium: the large grids)
      if constexpr (C >= 2 && MED == kDensity) {
        if (flags & 2)
          run_tile<T, C, D, 0, MED, 2>(a, lut_s, src, n0, tx, r0, r1);
        else
          run_tile<T, C, D, 0, MED, 1>(a, lut_s, src, n0, tx, r0, r1);
      } else {
        run_tile<T, C, D, 0, MED, 1>(a, lut_s, src, n0, tx, r0, r1);
      }
    } else {
      run_tile<T, C, D, 0, MED, 0>(a, lut_s, src, n0, tx, r0, r1);
    }
  }
}

// Persistent form for pools of independent shots: each work-group takes
// work-groups of 4 tile entries from the XCD range its index maps to (the
// lists are ordered shot by shot, so an XCD keeps to a few shots and their
// halo rows stay in its L2), then steals from the other ranges. The shot's
// StepArgs are read through the constant address space (scalar loads).

// MODE 0: interior tiles (list 0); 1: CPML-class tiles of every kind (list
// 1); 2: both, the CPML-class ones first (list 1, the fusable kinds; then
// list 0: one kernel for a whole launch, every wave takes CPML-class work
// while there is any, so the longer CPML-class tiles never run at the tail
// with the rest of the device idle); 3: the leftover CPML-class tiles (list
// 2) beside a MODE 2 kernel.
// The interior form runs at 4 waves per SIMD (at most 128 VGPRs; at 3 with
// two columns per lane and two rows of prefetch, which would spill 72 bytes
// per lane at 4); the forms with CPML-class tiles need more at T = 8 and run
// at 3 (a few VGPRs spill in some); the leftover form, which carries the
// general kind, at 2 rather than spill (its list is short).
template <int T, int C, int D, int MED, int MODE>
__global__ void __launch_bounds__(64 * kWavesPerWG, MODE == 0 ? (C > 1 && D > 1 ? 3 : 4) : MODE == 3 ? 2 : 3)
    pool_kernel(const StepArgs *__restrict__ descs, PoolQueue *q, int k, int tp) {
  __shared__ float4 lut_s[MED == kTable ? 256 : 1];
  __shared__ int s_pos, s_cls;
  const int g = __builtin_amdgcn_readfirstlane(q->g) + k;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int x = blockIdx.x & 7, tried = 0, cls_t = MODE == 0 ? 0 : MODE == 3 ? 2 : 1;  // used by thread 0 only
  int lut_slot = -1;
  for (;;) {
    if (threadIdx.x == 0) {
      int pos = -1;
      for (;;) {
        const int count = q->count[cls_t];
        while (tried < 8) {
          const int lo = static_cast<int>(static_cast<long long>(x) * count / 8);
          const int hi = static_cast<int>(static_cast<long long>(x + 1) * count / 8);
          const int t = atomicAdd(&q->ticket[k][cls_t][x], 1);
          if (lo + t < hi) {
            pos = lo + t;
            break;
          }
          x = (x + 1) & 7;
          ++tried;
        }
        if (pos >= 0 || MODE != 2 || cls_t == 0) break;
        cls_t = 0;  // the CPML-class list is drained: go on with the interior one
        x = blockIdx.x & 7;
        tried = 0;
      }
      s_pos = pos;
      s_cls = cls_t;
    }
    __syncthreads();
    const int pos = __builtin_amdgcn_readfirstlane(s_pos);
    const int cls = __builtin_amdgcn_readfirstlane(s_cls);
    if (pos < 0) return;  // every range is exhausted (uniform over the work-group)
    const int4 *items = q->items[cls];
    const long long e = 4LL * (4LL * pos + wave);
    const int slot = cld_i(items, 4LL * 4LL * pos + 3) >> 8;
    const StepArgs &a = *(const StepArgs *)((cargs_ptr)(descs + slot));
    if constexpr (MED == kTable) {
      if (slot != lut_slot) {
        for (int m = threadIdx.x; m < a.nmat; m += blockDim.x) lut_s[m] = gload4(a.lut, m);
        lut_slot = slot;
      }
      __syncthreads();
    }
    const int tx = cld_i(items, e), r0 = cld_i(items, e + 1), r1 = cld_i(items, e + 2), w = cld_i(items, e + 3);
    const int dg = a.g_stride == 1 ? g - a.g_base : (g - a.g_base) / a.g_stride;
    const int n0 = a.n_base + dg * tp;
    const int src = a.src_base ^ (dg & 1);
    if (r0 < r1) {
      if constexpr (MODE == 1 || MODE == 3) {
        run_cpml_tile<T, MED, MODE == 1 ? 0 : 2>(a, lut_s, src, n0, tx, r0, r1, w);
      } else {
        bool edge_tile = false;
        if constexpr (MODE == 2) {
          edge_tile = cls == 1;
          if (edge_tile) run_cpml_tile<T, MED, 1>(a, lut_s, src, n0, tx, r0, r1, w);
        }
        if (!edge_tile) {
          if (w & 1)
            run_tile<T, C, D, 0, MED, 1>(a, lut_s, src, n0, tx, r0, r1);
          else
            run_tile<T, C, D, 0, MED, 0>(a, lut_s, src, n0, tx, r0, r1);
        }
      }
    }
    // every wave is done with s_pos and the LUT before they are rewritten
    __syncthreads();
  }
}

// Launches of the pool kernel for one shape.
template <int T, int C, int D, int MODE>
hipError_t launch_pool_t(const StepArgs *descs, PoolQueue *q, int k, int tp, int med, int grid, hipStream_t s) {
  if (med == kTable)
    hipLaunchKernelGGL((pool_kernel<T, C, D, kTable, MODE>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, descs, q, k,
                       tp);
  else
    hipLaunchKernelGGL((pool_kernel<T, C, D, kDensity, MODE>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, descs, q,
                       k, tp);
  return hipGetLastError();
}

template <int C, int D, int MODE>
hipError_t launch_pool_d(const StepArgs *descs, PoolQueue *q, int k, int tp, int T, int med, int grid,
                         hipStream_t s) {
  switch (T) {
    case 1: return launch_pool_t<1, C, D, MODE>(descs, q, k, tp, med, grid, s);
    case 2: return launch_pool_t<2, C, D, MODE>(descs, q, k, tp, med, grid, s);
    case 3: return launch_pool_t<3, C, D, MODE>(descs, q, k, tp, med, grid, s);
    case 4: return launch_pool_t<4, C, D, MODE>(descs, q, k, tp, med, grid, s);
    case 5: return launch_pool_t<5, C, D, MODE>(descs, q, k, tp, med, grid, s);
    case 6: return launch_pool_t<6, C, D, MODE>(descs, q, k, tp, med, grid, s);
    case 7: return launch_pool_t<7, C, D, MODE>(descs, q, k, tp, med, grid, s);
    case 8: return launch_pool_t<8, C, D, MODE>(descs, q, k, tp, med, grid, s);
  }
  return hipErrorInvalidValue;
}

// Launches of the fused single-grid kernel (MODE 2) for one shape.
template <int T, int C, int D>
hipError_t launch_fused_tcd(const StepArgs &a, int src, int n0, hipStream_t s) {
  const int grid = a.n_wg[0];
  if (grid == 0) return hipSuccess;
  if (a.medium_kind == kTable)
    hipLaunchKernelGGL((step_kernel<T, C, D, kTable, 2>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, a, src, n0);
  else
    hipLaunchKernelGGL((step_kernel<T, C, D, kDensity, 2>), dim3(grid), dim3(64 * kWavesPerWG), 0, s, a, src, n0);
  return hipGetLastError();
}

template <int C, int D>
hipError_t launch_fused_cd(const StepArgs &a, int T, int src, int n0, hipStream_t s) {
  switch (T) {
    case 1: return launch_fused_tcd<1, C, D>(a, src, n0, s);
    case 2: return launch_fused_tcd<2, C, D>(a, src, n0, s);
    case 3: return launch_fused_tcd<3, C, D>(a, src, n0, s);
    case 4: return launch_fused_tcd<4, C, D>(a, src, n0, s);
    case 5: return launch_fused_tcd<5, C, D>(a, src, n0, s);
    case 6: return launch_fused_tcd<6, C, D>(a, src, n0, s);
    case 7: return launch_fused_tcd<7, C, D>(a, src, n0, s);
    case 8: return launch_fused_tcd<8, C, D>(a, src, n0, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace
}  // namespace fdtd2d

#endif
