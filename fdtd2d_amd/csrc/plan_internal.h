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
  int P;               // CPML thickness (0 = none)
  int xs;              // floats per row of the x-strip psi arrays (2*P rounded up)
  int nsteps;
  int medium_kind;
  int nmat;
  int pec;             // TABLE: material forced to Ez = 0 (-1 none)
  float db, ca0, cb0, cj, jz_decay, dxf, dyf;
  double jz_dt;
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
  // events by row
  const int32_t *ev_off;  // [ny+1]
  const Event *events;
  const float *src_table;
  int nsrc;
  float *det;
  int ndet;
  float *comp;
  int ncomp;
  long long comp_plane;
  // tiling
  int H;         // rows per band
  int ntx, nty;  // tiles across / down
  int nwgx;      // work-groups across (4 tiles each)
  const int32_t *tile_class;  // [nty * nwgx]: 1 if any tile of the WG touches a CPML strip
};

// One launch advances T steps (1..kMaxT).
constexpr int kMaxT = 4;
constexpr int kWavesPerWG = 4;

// Launches the step kernel for T steps on `stream`; src = buffer set read.
hipError_t launch_step(const StepArgs &a, int T, int C, int src, int n0, hipStream_t stream);

}  // namespace fdtd2d

#endif
