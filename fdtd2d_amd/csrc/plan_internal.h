// plan_internal.h - device-side layout shared by the stepper kernels and the
// plan code. HIP-only (included from .hip translation units).
#ifndef FDTD2D_PLAN_INTERNAL_H
#define FDTD2D_PLAN_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fdtd2d {

// Event kinds attached to a grid row, applied in this order at every step:
// soft source (Ez -= table[n][idx]), Ez detector, component detector.
enum : int { EV_SRC = 0, EV_DET = 1, EV_COMP = 2 };

struct Event {
  int32_t i;     // column
  int32_t kind;  // EV_*
  int32_t idx;   // element / detector / component index
  int32_t pad;
};

// Everything a step launch reads. Field buffers are ping-ponged: a launch of
// T steps reads set `src` and writes set 1-src, so that overlapping halo
// recomputation in neighbouring tiles always sees the old values.
struct StepArgs {
  int nx, ny, pitch;   // pitch: floats per row (multiple of 64), all fields
  int trash_rows;      // scratch rows below row ny-1 of every field plane
  int trash_slots;     // 64*C-float slots of those rows the waves spread their discarded stores over
  int P;               // CPML thickness (0 = none)
  int xs;              // floats per row of the x-strip psi arrays (2*P rounded up)
  int nsteps;
  int medium_kind;
  int nmat;
  int pec;             // TABLE: material forced to Ez = 0 (-1 none)
  float db, ca0, cb0, cj, jz_decay, dxf, dyf;
  double jz_dt;
  double rdx, rdy;  // the doubles nearest 1/dxf and 1/dyf
  // fields [2 buffers]
  float *ez[2], *hx[2], *hy[2], *jz[2];
  // psi [2 buffers]: y-strips [2P rows][pitch] by slot; x-strips [ny][xs] by slot
  float *psi_hxy[2], *psi_ezy[2], *psi_hyx[2], *psi_ezx[2];
  const int32_t *med_id;  // TABLE, pitch layout
  const float *med_ne;    // DENSITY, pitch layout
  const float4 *lut;      // TABLE: {ca, cb, ne, 0} per material
  // per-column tables [pitch], per-row tables [ny]
  const float *den_ex, *den_hx, *den_ey, *den_hy;
  const float *hyx_b, *hyx_c, *ezx_b, *ezx_c;  // [pitch]
  const int32_t *hyx_slot, *ezx_slot;          // [pitch], -1 outside the strips
  const float *hxy_b, *hxy_c, *ezy_b, *ezy_c;  // [ny]
  const int32_t *hxy_slot, *ezy_slot;          // [ny]
  // per row, kRowRec ints: {den_hy bits, den_ey bits, first event or -1 when
  // the row has none, end event, psi_hxy slot, psi_ezy slot (-1 outside the
  // y strips), hxy b, hxy c, ezy b, ezy c (bits), 0, 0}, for rows
  // -kMaxT .. ny+kMaxT-1 (the rows outside the grid: no events, no strips)
  const int *rowinfo;
  // events by row
  const int32_t *ev_off;  // [ny+1]
  const Event *events;
  const float *src_table;
  int nsrc;
  float *det;
  int ndet;
  // RC tracker capture: probe_col[pitch] = probe index of a column (-1: none);
  // the stepper stores Ez of every probe column for the step it computes into
  // rc_cap[((step-1) & 7)][probe][ny] (row = grid row), read after the launch.
  const int *probe_col;
  float *rc_cap;
  int nprobe;
  // Processing line: Ez of row line_j, columns [line_i0, line_i1), after the
  // sources of step n for line_begin < n <= line_end, into
  // line[(n - line_begin - 1) * (line_i1 - line_i0) + column - line_i0].
  // The row is special, so only the CPML-class kernel carries this code.
  float *line;
  int line_j, line_i0, line_i1, line_begin, line_end;
  float *comp;
  int ncomp;
  long long comp_plane;
  // den values of every row / column an interior tile loads (bitwise equal
  // there; rows and columns with other values make a tile CPML class)
  float uden_hx, uden_ex, uden_hy, uden_ey;
  // tiling (fixed per plan; launches of fewer steps reuse the same tiles)
  int vw;        // stored columns per tile = 64*C - 2*halo
  int halo;      // recomputed columns on each side, >= steps per launch
  int vw_e, halo_e;  // the same for CPML-class waves (1 column per lane)
  int x0;            // first column of interior tile 0
  int H;         // rows per band
  int ntx, nty;  // tiles across / down
  // Tile lists, one entry per wave = {tile, first row, end row, flags}; wave
  // w of work-group b takes entry 4b+w, entries with no rows are padding.
  // flags bit 0: the tile's window holds plasma or a nonzero Jz (interior
  // tiles without it skip the Jz update, load and store).
  // [0] interior tiles (no load touches a CPML cell), on bands of up to H
  // rows; [1] the rest, on sub-bands short enough to finish well inside the
  // interior launch they run beside. Each list is one launch.
  const int4 *tile_list[2];
  int n_tiles[2];  // entries in use
  int n_wg[2];     // work-groups launched (entries / 4, rounded up)
  // pool mode: the shot runs in launches g_base, g_base + g_stride, ...; at
  // its d-th such launch (d = (g - g_base) / g_stride) it runs steps
  // n_base + d*tp onwards from buffer set src_base ^ (d & 1)
  int g_base, n_base, src_base, g_stride;
};

// Work queue of the pool kernels, in device memory. items/count: the packed
// lists per class (work-groups of 4 tile entries); g: pool launches completed
// before the current chunk; ticket[k][cls][x]: work-groups taken from XCD
// range x of class cls by launch k of the chunk (reset by the tick kernel).
// items, count and g are written by kernels on the pool's stream, so every
// launch sees the lists in stream order, also inside a replayed graph.
constexpr int kMaxChunk = 16;
// list classes: 0 interior, 1 CPML class (all of it, or in a fused pool the
// fusable kinds), 2 the leftover CPML-class entries of a fused pool
constexpr int kPoolLists = 3;
struct PoolQueue {
  const int4 *items[kPoolLists];
  int count[kPoolLists];
  int g;
  int pad;
  int ticket[kMaxChunk][kPoolLists][8];
};

constexpr int kRowRec = 12;  // ints per row record

// One launch advances T steps (1..kMaxT).
constexpr int kMaxT = 8;
// CPML-class list entries run one column per lane whatever the interior
// kernel's columns per lane: their first field packs the first stored column
// (low kEdgeColBits bits) and the stored width (above).
constexpr int kEdgeColBits = 20;
constexpr int kWavesPerWG = 4;
// Kind of a CPML-class entry (flags bits 2-4), from the window the entry
// loads. Rows clear: no y-strip, ring or special-den row; columns clear: no
// x-strip, ring or special-den column; events: a source, detector or
// component cell, an RC probe column or the processing line's row.
//   x strip: no events, rows clear         y strip: no events, columns clear
//   corner:  no events                     events: rows and columns clear
//   general: the rest (an event inside a strip or beside a special row)
enum : int { kKindGeneral = 0, kKindX = 1, kKindY = 2, kKindXY = 3, kKindE = 4 };
constexpr int kKindShift = 2;
inline __host__ __device__ int entry_kind(int flags) { return (flags >> kKindShift) & 7; }
// A CPML-class entry in the one list of a fused single-grid launch.
constexpr int kFlagCpml = 32;
constexpr int kAllKinds = 31;
// CPML-class kinds the fused single-grid kernel carries (bit 1 << kind) for
// a medium (0 table, 1 density): those that keep it within 3 waves per SIMD
// with no scratch spills at 8 steps per launch. The density medium's kernel
// carries every kind (165 VGPRs); the table medium's only the strips and
// the event kind without captures (with the corner kind it spills, and the
// capturing row loop of event entries of problems with RC probes or a
// processing line would too); its other entries run in a kernel beside it.
constexpr int fused_kinds(int medium_kind) {
  return medium_kind == 1 ? kAllKinds : (1 << kKindX) | (1 << kKindY) | (1 << kKindE);
}
// Whether a CPML-class entry of this kind goes to the fused kernel (else to
// the leftover kernel beside it), for a problem with or without captures
// (RC probes, a processing line).
constexpr bool fusable_kind(int kind, int medium_kind, bool captures) {
  return fused_kinds(medium_kind) == kAllKinds ||
         (((fused_kinds(medium_kind) >> kind) & 1) && !(kind == kKindE && captures));
}

// Launches the step kernel of work-group class `cls` (0 interior, 1 CPML,
// every kind; 2 CPML, the leftover kinds beside a fused launch) for T steps
// on `stream`, D rows of prefetch; src = buffer set read.
hipError_t launch_step(const StepArgs &a, int cls, int T, int C, int D, int src, int n0, hipStream_t stream);
// Launches the fused single-grid kernel over the combined list (tile_list[0],
// n_wg[0] work-groups: CPML-class entries flagged kFlagCpml first in each
// XCD's range, then interior ones). C = 1 or 2.
hipError_t launch_step_fused(const StepArgs &a, int T, int C, int D, int src, int n0, hipStream_t stream);

// RC cutoff lines (Dev_RCcalc semantics) for steps n0+1 .. n0+t from the
// capture of the launch that computed them; one work-group per line.
struct RcArgs {
  const float *cap;
  const int *map;  // [2*nrc]: probe of the maximum column, probe of the scan column
  float *edge;     // [nrc] edge levels
  int *spd;        // [(nsteps+1)*nrc]
  int nrc, nprobe, ny, jcut;
  float level;
};
hipError_t launch_rc(const RcArgs &r, int n0, int t, hipStream_t stream);

// Pool kernels: launch k of a chunk, T steps, on `grid` persistent
// work-groups that take work-groups of tile entries (entry.w = flags | shot
// slot << 8, one slot per work-group) from q's lists until they are
// exhausted. cls: 0 interior (list 0), 1 CPML class (list 1, every kind),
// 2 fused (list 1 of fusable kinds, then list 0), 3 the leftover CPML-class
// entries (list 2). tp = the pool's steps per full launch. C = 1 or 2
// columns per lane and D = 1 or 2 rows of prefetch for the interior tiles of
// classes 0 and 2 (CPML-class tiles run at one column per lane).
hipError_t launch_pool(const StepArgs *descs, PoolQueue *q, int k, int tp, int cls, int T, int C, int D,
                       int medium_kind, int grid, hipStream_t stream);
// The C = 2 forms of classes 0 and 2 (their own translation unit).
hipError_t launch_pool_c2(const StepArgs *descs, PoolQueue *q, int k, int tp, int cls, int T, int D,
                          int medium_kind, int grid, hipStream_t stream);
// Resets the tickets of launches [0, nk) and adds dg to q->g (one wave).
hipError_t launch_pool_tick(PoolQueue *q, int nk, int dg, hipStream_t stream);
// Points q at new lists (g >= 0 also sets q->g), in stream order.
hipError_t launch_pool_set(PoolQueue *q, const int4 *const items[kPoolLists], const int count[kPoolLists], int g,
                           hipStream_t stream);

// True when tile tx (columns from x0 + tx*vw, vw stored, halo recomputed
// columns each side) over output rows [jb0, jb1) loads a cell of a CPML
// strip or of the outer ring (with no CPML the ring counts as a strip of
// width 1): such a tile takes the CPML-class kernel.
inline bool tile_touches_cpml(int nx, int ny, int P, int x0, int vw, int halo, int T, int tx, int jb0, int jb1) {
  const int Pe = P > 0 ? P : 1;
  const int jl0 = jb0 - T > 0 ? jb0 - T : 0, jl1 = jb1 + T < ny ? jb1 + T : ny;
  if (jl0 <= Pe || jl1 >= ny - Pe - 1) return true;
  const int cs = x0 + tx * vw - halo, ce = cs + vw + 2 * halo;
  return cs <= Pe || ce >= nx - Pe - 1;
}

}  // namespace fdtd2d

#endif
