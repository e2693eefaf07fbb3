// plan_impl.h - internals shared by plan.hip (single plans) and pool.hip
// (pools of independent shots advanced by the persistent kernels).
#ifndef FDTD2D_PLAN_IMPL_H
#define FDTD2D_PLAN_IMPL_H

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "../../include/fdtd2d.h"
#include "plan_internal.h"

namespace fdtd2d {
namespace detail {

constexpr size_t kAlign = 256;
// Scratch rows below each field plane: interior tiles direct the stores of
// lanes/rows they do not own there, so their loop issues the same memory
// operations every row (nothing reads these rows).
constexpr int kTrashRows = 64;
// Discarded stores (halo lanes, rows outside the band) go to this many
// 64*C-float slots: few enough to stay resident in L2, so they cost no HBM
// writes, and spread over enough lines to avoid one hot L2 channel.
constexpr int kTrashSlots = 64;
// Steps per launch when none is set: on MI355X the interior kernel is issue
// bound near 8 levels and memory bound below (8192^2: 81 Gcell/s at 1,
// ~300 at 4, ~520 at 8 with 1 column per lane and 128-row bands).
constexpr int kDefaultT = 8;

struct Layout {
  int nx, ny, pitch, P, xs;
  size_t off_fields[2][4];
  size_t off_psi_y[2][2];  // [set][hxy, ezy]
  size_t off_psi_x[2][2];  // [set][hyx, ezx]
  size_t off_med, off_lut;
  size_t off_col[8];  // den_ex, den_hx, hyx_b, hyx_c, ezx_b, ezx_c, hyx_slot, ezx_slot
  size_t off_row[8];  // den_ey, den_hy, hxy_b, hxy_c, ezy_b, ezy_c, hxy_slot, ezy_slot
  size_t off_rowinfo, off_evoff, off_events, off_src, off_det, off_comp;
  size_t off_probe, off_cap, off_rcmap, off_rcedge, off_spd;  // RC tracker
  size_t off_line;  // processing line captures
  size_t line_rows;
  int nprobe;  // distinct columns the RC lines read
  size_t n_events;
  size_t total;
};

// Tile lists of one launch shape, one entry per wave: [0] interior
// {tx, r0, r1, flags}, [1] CPML class {edge_entry_x(c0, w), r0, r1, flags}.
// A (tile, band) whose loads touch a CPML strip or the outer ring is cut into
// sub-bands of hc rows; the sub-bands that still touch go to the CPML list,
// split into pieces of vw_e columns, runs of the others are merged back into
// one interior entry. Host-only, so the partition can be checked without a
// GPU.
inline int edge_entry_x(int c0, int w) { return c0 | (w << kEdgeColBits); }
inline int edge_entry_c0(int x) { return x & ((1 << kEdgeColBits) - 1); }
inline int edge_entry_w(int x) { return x >> kEdgeColBits; }

struct Tiling {
  int T = 0, halo = 0, vw = 0, ntx = 0, H = 0, nty = 0, hc = 0;
  int halo_e = 0, vw_e = 0;  // CPML-class waves: 1 column per lane
  // Interior tile tx stores columns from x0 + tx*vw. With more than one
  // column per lane the interior tiles start after a left margin of CPML-class
  // pieces wide enough for the x-strip and end before a right margin of
  // width >= x0; with one column per lane x0 = 0 and the tiles cover all.
  int x0 = 0;
  std::vector<int4> lists[2];  // padded to whole work-groups
  int n_tiles[2] = {0, 0};
  int n_nojz = 0;  // interior tiles that skip the Jz update
  int n_full = 0;  // interior tiles whose window is full of plasma
  double cells[2] = {0, 0};
  double cells_kind[5] = {0, 0, 0, 0, 0};  // CPML-class cells by kind (kKindGeneral ... kKindE)
};

struct TileMasks {
  // prefix counts of the rows / columns whose den differ from the uniform
  // interior values (and the last row / column): [ny+1], [nx+1]
  const int *srow_pref = nullptr, *scol_pref = nullptr;
  // prefix counts of the processing line's row and of the RC probe columns
  const int *lrow_pref = nullptr, *pcol_pref = nullptr;
  const int *ne_lo = nullptr, *ne_hi = nullptr, *jz_lo = nullptr, *jz_hi = nullptr;  // [ny]
  // per row, the runs [x, y) of consecutive columns with Ne > 0: row r's
  // runs are ne_runs[ne_run_off[r] .. ne_run_off[r+1])
  const int *ne_run_off = nullptr;  // [ny+1]
  const int2 *ne_runs = nullptr;
  const int2 *ev = nullptr;  // event cells {column, row}, sorted by (row, column)
  size_t nev = 0;
};

Tiling make_tiling(int nx, int ny, int P, int T, int C, int band, int band_cpml, const TileMasks *m = nullptr);
Layout make_layout(const fdtd2d_problem *p, size_t n_events);
int hip_fail(hipError_t e, const char *what);

}  // namespace detail
}  // namespace fdtd2d

struct fdtd2d_plan {
  bool pooled = false;  // a pool shot: no streams of its own, host tile lists only
  bool owns_base = true;  // false: base is carved from a pool's arena or given by the caller
  // created by fdtd2d_plan_create_in: tile lists in the library's static
  // per-device array, no streams or events of its own
  bool static_lists = false;
  bool holds_static = false;  // the plan owns its device's static array (released by destroy)
  fdtd2d::detail::Tiling tiling;  // pooled: the tile lists the pool concatenates
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  char *base = nullptr;
  fdtd2d::detail::Layout L{};
  fdtd2d::StepArgs a{};
  int src = 0;
  int n = 0;
  int nsteps = 0;
  int T = 0;  // 0: default
  int C = 1;
  int band = 0;  // 0: automatic
  int D = 2;     // rows of prefetch per wave (interior kernel); the row loop is unrolled by D
  int band_cpml = 0;               // rows per CPML-class work-group (0: automatic)
  int jz_first = 1;  // interior work-groups with a Jz update first in each XCD's range
  int fused = 1;     // one kernel per launch over one list (0: interior and CPML-class kernels on two streams)
  double cells_class[2] = {0, 0};  // grid cells owned by interior / CPML-class tiles
  double cells_kind[5] = {0, 0, 0, 0, 0};  // CPML-class cells by kind: general, x strip, y strip, corner, events
  int hc = 0;                      // rows per CPML-class sub-band in use
  int n_nojz = 0;                  // interior tiles skipping the Jz update
  int n_full = 0;                  // interior tiles whose window is full of plasma
  int4 *wg_buf = nullptr;          // device copy of the two work-group lists
  hipStream_t aux = nullptr;       // CPML-class launches run here, beside the interior ones
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int ndet = 0, ncomp = 0, nsrc = 0;
  bool timing = false;
  // timed launches per kernel class (0 interior on the plan's stream, 1
  // CPML class on the aux stream): event pairs not yet read, and the totals
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[2];
  std::vector<hipEvent_t> spare;
  double sum_ms[2] = {0, 0};
  int count[2] = {0, 0};
  int launches_T[fdtd2d::kMaxT + 1] = {0};
  // Host facts the tiling needs. Prefix counts of the rows / columns whose
  // den values differ from the uniform interior ones (the interior kernel
  // takes those as scalars), and per row the column span [lo, hi] holding
  // plasma (Ne > 0) or a nonzero Jz: windows outside every span skip the Jz
  // update (there Jz stays +0 and Ez - Jz = Ez).
  std::vector<int> srow_pref, scol_pref;
  std::vector<int> lrow_pref, pcol_pref;  // the processing line's row, the RC probe columns
  std::vector<int> ne_lo, ne_hi, jz_lo, jz_hi;
  // runs of plasma columns per row (see TileMasks): an interior tile whose
  // whole window lies inside them skips the Ne > 0 test (flag bit 1)
  std::vector<int> ne_run_off;
  std::vector<int2> ne_runs;
  // cells holding a source, detector or probe, sorted by (row, column): the
  // tiles whose window holds one take the CPML-class kernel, so the interior
  // kernel carries no event code
  std::vector<int2> ev_cells;
  size_t line_rows = 0;  // processing line: captured steps
  int line_size = 0;      // and cells per step
  int nrc = 0;          // RC cutoff lines
  fdtd2d::RcArgs rc{};  // their device state
};

namespace fdtd2d {
namespace detail {

// Creates a plan. pooled: no streams or events of its own and tile lists
// kept on the host for a pool, whose launch shape (T, C, band, band_cpml) it
// takes; work goes to `stream`.
// base: device memory of at least make_layout(p, events).total bytes to use
// instead of a hipMalloc of its own (nullptr: allocate).
// caller_mem: base is the caller's device memory and stream the caller's
// stream; the plan creates no stream or event and allocates no device
// memory (its tile lists go to the static array, fdtd2d_plan_create_in).
int create_plan(const fdtd2d_problem *p, int device, bool pooled, hipStream_t stream, int T, int C, int band,
                int band_cpml, char *base, fdtd2d_plan **out, bool caller_mem = false);
// Bytes create_plan needs for p (the plan's own allocation).
size_t plan_bytes(const fdtd2d_problem *p);
int set_tiling(fdtd2d_plan *pl);
int current_T(const fdtd2d_plan *pl);

}  // namespace detail
}  // namespace fdtd2d

#define HIP_TRY(expr, what)                                            \
  do {                                                                 \
    hipError_t e_ = (expr);                                            \
    if (e_ != hipSuccess) return ::fdtd2d::detail::hip_fail(e_, what); \
  } while (0)

#endif
