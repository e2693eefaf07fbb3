// plan.hip - device plans: the MI355X replacement of DevMemoryAllocation,
// the DevFieldCalculation step loop and DevMemoryFree
// (F/kernel.cu:1087-1618, 2065-2309, 2820-2987).
//
// HBM layout of one plan (one hipMalloc, carved at 256-byte boundaries):
//   fields   Ez, Hx, Hy, Jz     2 buffer sets x ny x pitch floats (pitch =
//                               nx rounded up to 64; padding stays 0)
//   psi      y strips           2 sets x {Hxy, Ezy} x 2P rows x pitch
//            x strips           2 sets x {Hyx, Ezx} x ny x xs (xs = 2P)
//   medium   ny x pitch          int32 material id (TABLE) or float Ne (DENSITY)
//   tables   per column [pitch], per row [ny], material LUT, row events,
//            source table [(nsteps+1) x nsrc], traces [(nsteps+1) x ndet],
//            components [3 x (nsteps+1) x ncomp]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/fdtd2d.h"
#include "errors.h"
#include "plan_impl.h"
#include "plan_internal.h"
#include "runs.h"
#include "validate.h"

using fdtd2d::set_error;
using namespace fdtd2d::detail;

namespace fdtd2d {
namespace detail {


size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }


std::vector<fdtd2d::Event> build_events(const fdtd2d_problem *p, std::vector<int32_t> &off) {
  struct E {
    int j;
    fdtd2d::Event e;
  };
  std::vector<E> v;
  for (int e = 0; e < p->nsrc; ++e) v.push_back({p->src_j[e], {p->src_i[e], fdtd2d::EV_SRC, e, 0}});
  for (int d = 0; d < p->ndet; ++d) v.push_back({p->det_j[d], {p->det_i[d], fdtd2d::EV_DET, d, 0}});
  for (int c = 0; c < p->ncomp; ++c) v.push_back({p->comp_j[c], {p->comp_i[c], fdtd2d::EV_COMP, c, 0}});
  std::stable_sort(v.begin(), v.end(), [](const E &a, const E &b) {
    if (a.j != b.j) return a.j < b.j;
    if (a.e.kind != b.e.kind) return a.e.kind < b.e.kind;
    return a.e.idx < b.e.idx;
  });
  off.assign(p->ny + 1, 0);
  for (auto &x : v) off[x.j + 1]++;
  for (int j = 0; j < p->ny; ++j) off[j + 1] += off[j];
  std::vector<fdtd2d::Event> out;
  for (auto &x : v) out.push_back(x.e);
  return out;
}

Layout make_layout(const fdtd2d_problem *p, size_t n_events) {
  Layout L{};
  L.nx = p->nx;
  L.ny = p->ny;
  L.pitch = static_cast<int>(align_up(p->nx, 64));
  L.P = p->npml;
  L.xs = std::max(1, 2 * p->npml);
  L.n_events = n_events;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t r = o;
    o = align_up(o + std::max<size_t>(bytes, 4), kAlign);
    return r;
  };
  const size_t plane = static_cast<size_t>(L.ny) * L.pitch * 4;
  const size_t fplane = static_cast<size_t>(L.ny + kTrashRows) * L.pitch * 4;
  for (int s = 0; s < 2; ++s)
    for (int f = 0; f < 4; ++f) L.off_fields[s][f] = take(fplane);
  const size_t ystrip = static_cast<size_t>(2 * std::max(L.P, 1)) * L.pitch * 4;
  const size_t xstrip = static_cast<size_t>(L.ny) * L.xs * 4;
  for (int s = 0; s < 2; ++s)
    for (int f = 0; f < 2; ++f) L.off_psi_y[s][f] = take(ystrip);
  for (int s = 0; s < 2; ++s)
    for (int f = 0; f < 2; ++f) L.off_psi_x[s][f] = take(xstrip);
  L.off_med = take(plane);
  L.off_lut = take(256 * 16);
  for (int f = 0; f < 8; ++f) L.off_col[f] = take(static_cast<size_t>(L.pitch) * 4);
  for (int f = 0; f < 8; ++f) L.off_row[f] = take(static_cast<size_t>(L.ny) * 4);
  L.off_rowinfo = take(static_cast<size_t>(L.ny + 2 * fdtd2d::kMaxT) * fdtd2d::kRowRec * 4);
  L.off_evoff = take(static_cast<size_t>(L.ny + 1) * 4);
  L.off_events = take(n_events * sizeof(fdtd2d::Event));
  L.off_src = take(static_cast<size_t>(p->nsteps + 1) * p->nsrc * 4);
  L.off_det = take(static_cast<size_t>(p->nsteps + 1) * p->ndet * 4);
  L.off_comp = take(static_cast<size_t>(3) * (p->nsteps + 1) * p->ncomp * 4);
  {
    std::vector<int> cols;
    for (int e = 0; e < p->nrc; ++e) {
      cols.push_back(p->rc_i[e]);
      cols.push_back(p->rc_imax[e]);
    }
    std::sort(cols.begin(), cols.end());
    L.nprobe = static_cast<int>(std::unique(cols.begin(), cols.end()) - cols.begin());
  }
  L.off_probe = take(static_cast<size_t>(L.pitch) * 4);
  L.off_cap = take(static_cast<size_t>(fdtd2d::kMaxT) * L.nprobe * L.ny * 4);
  L.off_rcmap = take(static_cast<size_t>(2) * p->nrc * 4);
  L.off_rcedge = take(static_cast<size_t>(p->nrc) * 4);
  L.off_spd = take(static_cast<size_t>(p->nsteps + 1) * p->nrc * 4);
  L.line_rows =
      p->line_size > 0 ? static_cast<size_t>(std::max(0, std::min(p->line_end, p->nsteps) - p->line_begin)) : 0;
  L.off_line = take(L.line_rows * std::max(p->line_size, 0) * 4);
  L.total = o;
  return L;
}

}  // namespace detail
}  // namespace fdtd2d


namespace fdtd2d {
namespace detail {

int hip_fail(hipError_t e, const char *what) {
  return set_error(FDTD2D_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------- static lists
// The tile lists (the step kernels' work lists) of a plan made by
// fdtd2d_plan_create_in live in this array instead of a hipMalloc of their
// own: kStaticListEntries entries of 16 bytes (1 MiB) on every device the
// library runs on, held until the process ends. Host code never passes the
// array's name to a kernel or a copy (in host code it is a stand-in, not a
// device address): it takes the device's own address from
// hipGetSymbolAddress once per device, with that device current.
constexpr size_t kStaticListEntries = 65536;
static __device__ __align__(256) int4 scratch[kStaticListEntries];

constexpr int kMaxDevices = 64;
static std::mutex g_static_mu;
static int4 *g_static_ptr[kMaxDevices];
static bool g_static_busy[kMaxDevices];

// The device address of `scratch` on `device` (which must be current).
static int static_lists_ptr(int device, int4 **ptr) {
  std::lock_guard<std::mutex> lk(g_static_mu);
  if (!g_static_ptr[device]) {
    void *q = nullptr;
    hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !q) return hip_fail(e, "hipGetSymbolAddress of the static tile-list array");
    g_static_ptr[device] = static_cast<int4 *>(q);
  }
  *ptr = g_static_ptr[device];
  return FDTD2D_OK;
}

// Zeroes n 16-byte words (a kernel, not hipMemset: callers may capture the
// stream the plan works on).
__global__ void __launch_bounds__(256) zero_words(uint4 *p, size_t n) {
  for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}


// Band height: enough bands for >= ~8 waves per SIMD on 256 CUs, at least 32
// rows, at most 256. With two columns per lane (a wider tile, about twice
// the state per wave) about 6 waves per SIMD, from a finer set of heights
// up to 192: 8192^2 then takes 96-row bands (measured 752 Gcell/s against
// 703 at 128 and 735 at 64), 16384^2 192-row bands (800 against 744 at 256,
// 757 at 128 and 781 at 96).
int pick_band(int nx, int ny, int VW, int T, int C) {
  const int ntx = (nx + VW - 1) / VW;
  auto waves = [&](int H) { return static_cast<long long>(ntx) * ((ny + H - 1) / H); };
  if (C >= 2) {
    // the tallest band that still gives about 6 waves per SIMD (16384^2: 192
    // rows, 8192^2: 96); smaller grids take 64 (4096^2: 725 Gcell/s against
    // 606 for one column per lane and 32-row bands)
    static const int heights[] = {192, 160, 128, 112, 96, 80, 64};
    for (int H : heights)
      if (waves(H) >= 6LL * 1024) return std::max(H, T);
    return std::max(64, T);
  }
  int H = 256;
  while (H > 32 && waves(H) < 8LL * 1024) H /= 2;
  return std::max(H, T);
}

// Columns per lane a plan starts with: two when the grid gives every SIMD
// about 2 such waves at 64-row bands (8192^2: 752 against 713 Gcell/s;
// 4096^2: 725 against 606-616), else one (2048^2: one column per lane 324
// against 295).
int default_cols_per_lane(int nx, int ny) {
  const long long ntx2 = (nx + 111) / 112;  // 64*2 - 2*8 stored columns
  return ntx2 * (ny / 64) >= 2048 ? 2 : 1;
}


Tiling make_tiling(int nx, int ny, int P, int T, int C, int band, int band_cpml, const TileMasks *m) {
  Tiling t;
  t.T = T;
  t.halo = ((T + C - 1) / C) * C;
  t.vw = 64 * C - 2 * t.halo;
  t.halo_e = T;
  t.vw_e = 64 - 2 * t.halo_e;
  const int Pe = std::max(P, 1);
  // A CPML-class piece at the grid's left or right edge needs no halo on
  // that side (the edge cells depend on nothing beyond the grid): it stores
  // up to we columns of the 64 it loads.
  const int we = 64 - t.halo_e;
  // Interior tile tx stores [x0 + tx*vw, x0 + (tx+1)*vw); its window (halo
  // columns on each side) must load no column of the x strips or the ring:
  // x0 - halo > Pe, and the last tile's window ends before nx - Pe - 1. The
  // margins [0, x0) and [xr, nx) go to CPML-class pieces. x0 is a multiple
  // of 16 floats, so that the tiles' stored rows start on 64-byte
  // boundaries (vw is a multiple of 16 at 8 steps per launch): at 8192^2 the
  // interior kernel alone measured 0.667 ms per launch with x0 = 56 (rows
  // starting 32 bytes into a 64-byte segment) against 0.608 with x0 = 48.
  // Of the x0 that keep the left margin one piece (at most we columns),
  // the one that leaves the fewest right-margin pieces (a piece costs a
  // whole wave per row, whatever its width); ties to the smallest x0.
  auto right_pieces = [&](int x0, int ntx) {
    const int m = ntx > 0 ? nx - (x0 + ntx * t.vw) : nx;
    return m <= we ? 1 : 1 + (m - we + t.vw_e - 1) / t.vw_e;
  };
  t.x0 = (Pe + t.halo + 1 + 15) / 16 * 16;
  t.ntx = std::max(0, (nx - Pe - 2 - t.halo - t.x0) / t.vw);
  for (int x0 = t.x0 + 16; x0 <= we && t.ntx > 0; x0 += 16) {
    const int ntx = std::max(0, (nx - Pe - 2 - t.halo - x0) / t.vw);
    if (ntx == t.ntx && right_pieces(x0, ntx) < right_pieces(t.x0, t.ntx)) t.x0 = x0;
  }
  const int xr = t.ntx > 0 ? t.x0 + t.ntx * t.vw : 0;  // right margin [xr, nx)
  t.H = band > 0 ? std::max(band, T) : pick_band(nx, ny, t.vw, T, C);
  t.nty = (ny + t.H - 1) / t.H;
  t.hc = band_cpml > 0 ? std::max(band_cpml, T) : std::max(32, T);
  auto count = [](const int *pref, int a, int b) { return pref ? pref[b] - pref[a] : 0; };
  // an event cell (source, detector, component detector) in rows [jl0, jl1), columns [cs, ce)
  auto events_in = [&](int jl0, int jl1, int cs, int ce) {
    if (!m) return false;
    size_t lo = 0, hi = m->nev;  // first event at or after row jl0
    while (lo < hi) {
      const size_t mid = (lo + hi) / 2;
      if (m->ev[mid].y < jl0) lo = mid + 1;
      else hi = mid;
    }
    for (size_t q = lo; q < m->nev && m->ev[q].y < jl1; ++q)
      if (m->ev[q].x >= cs && m->ev[q].x < ce) return true;
    return false;
  };
  // window of rows / columns an interior tile loads
  auto window = [&](int tx, int r0, int r1, int &jl0, int &jl1, int &cs, int &ce) {
    jl0 = std::max(0, r0 - T);
    jl1 = std::min(ny, r1 + T);
    cs = std::max(0, t.x0 + tx * t.vw - t.halo);
    ce = std::min(nx, t.x0 + tx * t.vw + t.vw + t.halo);
  };
  auto touches = [&](int tx, int r0, int r1) {
    if (fdtd2d::tile_touches_cpml(nx, ny, P, t.x0, t.vw, t.halo, T, tx, r0, r1)) return true;
    if (!m) return false;
    int jl0, jl1, cs, ce;
    window(tx, r0, r1, jl0, jl1, cs, ce);
    if (count(m->srow_pref, jl0, jl1) || count(m->lrow_pref, jl0, jl1) || count(m->scol_pref, cs, ce) ||
        count(m->pcol_pref, cs, ce))
      return true;
    return events_in(jl0, jl1, cs, ce);
  };
  // Kind of a CPML-class piece storing columns [c0, c1) of rows [r0, r1),
  // from the window it loads (see kKindGeneral .. kKindE), as flag bits.
  auto piece_kind = [&](int c0, int r0, int r1) {
    const int cs = std::max(0, c0 - t.halo_e), ce = std::min(nx, cs + 64);
    const int jl0 = std::max(0, r0 - T), jl1 = std::min(ny, r1 + T);
    const bool ev =
        m && (events_in(jl0, jl1, cs, ce) || count(m->lrow_pref, jl0, jl1) || count(m->pcol_pref, cs, ce));
    const bool rows = jl0 > Pe && jl1 < ny - Pe - 1 && !(m && count(m->srow_pref, jl0, jl1));
    const bool cols = cs > Pe && ce < nx - Pe - 1 && !(m && count(m->scol_pref, cs, ce));
    int kind = fdtd2d::kKindGeneral;
    if (!ev) kind = rows ? fdtd2d::kKindX : cols ? fdtd2d::kKindY : fdtd2d::kKindXY;
    else if (rows && cols) kind = fdtd2d::kKindE;
    return kind << fdtd2d::kKindShift;
  };
  const int kX_flags = fdtd2d::kKindX << fdtd2d::kKindShift;
  std::vector<int4> *lists = t.lists;
  auto add_piece = [&](int c0, int c1, int r0, int r1, int kind) {
    lists[1].push_back(make_int4(edge_entry_x(c0, c1 - c0), r0, r1, 1 | kind));
  };
  // rows [ytop, ybot) are where a band's window (T rows of cone above and
  // below) can stay clear of the y strips and the ring
  const int ytop = Pe + T + 1, ybot = ny - Pe - T - 2;
  // Sub-bands of band [jb0, jb1): hc rows, ending at the first row whose
  // window clears the top strip and at the last one clear of the bottom
  // strip, so the rows between them and the strips can run as longer tiles
  // (when the band keeps at least a sub-band's height of them: shorter ones
  // cost more in cone rows than they save).
  auto sub_bands = [&](int jb0, int jb1, std::vector<int2> &out) {
    out.clear();
    for (int r0 = jb0, r1; r0 < jb1; r0 = r1) {
      r1 = std::min(jb1, r0 + t.hc);
      if (r0 < ytop && r1 > ytop && jb1 - ytop >= t.hc) r1 = ytop;
      if (r0 < ybot && r1 > ybot && ybot - jb0 >= t.hc) r1 = ybot;
      out.push_back(make_int2(r0, r1));
    }
  };
  std::vector<int2> subs;
  for (int tx = 0; tx < t.ntx; ++tx)
    for (int ty = 0; ty < t.nty; ++ty) {
      const int jb0 = ty * t.H, jb1 = std::min(ny, jb0 + t.H);
      if (!touches(tx, jb0, jb1)) {
        lists[0].push_back(make_int4(tx, jb0, jb1, 0));
        continue;
      }
      sub_bands(jb0, jb1, subs);
      int run0 = -1;  // start of the current run of interior sub-bands
      for (const int2 &sb : subs) {
        const int r0 = sb.x, r1 = sb.y;
        if (touches(tx, r0, r1)) {
          if (run0 >= 0) lists[0].push_back(make_int4(tx, run0, r0, 0));
          run0 = -1;
          const int c1 = std::min(nx, t.x0 + tx * t.vw + t.vw);
          for (int c0 = t.x0 + tx * t.vw; c0 < c1; c0 += t.vw_e) {
            const int e1 = std::min(c0 + t.vw_e, c1);
            add_piece(c0, e1, r0, r1, piece_kind(c0, r0, r1));
          }
        } else if (run0 < 0) {
          run0 = r0;
        }
      }
      if (run0 >= 0) lists[0].push_back(make_int4(tx, run0, jb1, 0));
    }
  // The margins: pieces of one column per lane across [0, x0) and [xr, nx)
  // (all of [0, nx) when no interior tile fits), on whole bands where the
  // band's window makes them x-strip pieces, else on sub-bands, each of its
  // own kind, with runs of x-strip sub-bands merged.
  std::vector<int2> cols;
  auto margin_cols = [&](int a, int b) {
    for (int c = a; c < b;) {
      int w;
      if (c == 0) w = (b == nx && nx <= 64) ? nx : std::min(b, we);
      else if (b == nx && c >= nx - we) w = nx - c;
      else if (b == nx) w = std::min(t.vw_e, nx - we - c);
      else w = std::min(t.vw_e, b - c);
      cols.push_back(make_int2(c, c + w));
      c += w;
    }
  };
  if (t.ntx > 0) {
    margin_cols(0, t.x0);
    margin_cols(xr, nx);
  } else {
    margin_cols(0, nx);
  }
  for (const int2 &pc : cols)
    for (int ty = 0; ty < t.nty; ++ty) {
      const int jb0 = ty * t.H, jb1 = std::min(ny, jb0 + t.H);
      if (piece_kind(pc.x, jb0, jb1) == kX_flags) {
        add_piece(pc.x, pc.y, jb0, jb1, kX_flags);
        continue;
      }
      sub_bands(jb0, jb1, subs);
      int run0 = -1;  // start of the current run of x-strip sub-bands
      for (const int2 &sb : subs) {
        const int kind = piece_kind(pc.x, sb.x, sb.y);
        if (kind == kX_flags) {
          if (run0 < 0) run0 = sb.x;
          continue;
        }
        if (run0 >= 0) add_piece(pc.x, pc.y, run0, sb.x, kX_flags);
        run0 = -1;
        add_piece(pc.x, pc.y, sb.x, sb.y, kind);
      }
      if (run0 >= 0) add_piece(pc.x, pc.y, run0, jb1, kX_flags);
    }
  auto needs_jz = [&](const int4 &w) {
    if (!m || !m->ne_lo) return 1;
    int jl0, jl1, cs, ce;
    window(w.x, w.y, w.z, jl0, jl1, cs, ce);
    for (int r = jl0; r < jl1; ++r) {
      if (m->ne_lo[r] <= m->ne_hi[r] && m->ne_lo[r] < ce && m->ne_hi[r] >= cs) return 1;
      if (m->jz_lo[r] <= m->jz_hi[r] && m->jz_lo[r] < ce && m->jz_hi[r] >= cs) return 1;
    }
    return 0;
  };
  // every cell of the window has Ne > 0: each of its rows has a plasma run
  // covering the window's columns
  auto plasma_full = [&](const int4 &w) {
    if (!m || !m->ne_run_off) return false;
    int jl0, jl1, cs, ce;
    window(w.x, w.y, w.z, jl0, jl1, cs, ce);
    for (int r = jl0; r < jl1; ++r) {
      bool in = false;
      for (int q = m->ne_run_off[r]; q < m->ne_run_off[r + 1] && !in; ++q)
        in = m->ne_runs[q].x <= cs && m->ne_runs[q].y >= ce;
      if (!in) return false;
    }
    return true;
  };
  // Order by column group of 4 tiles, then first row, then tile: the 4 waves
  // of a work-group get neighbouring tiles and consecutive work-groups go down
  // the same columns. Pad each list to whole work-groups with empty entries.
  for (int c = 0; c < 2; ++c) {
    // column-group index of an entry: tile / 4, or first column / (4 vw_e)
    auto group = [&](const int4 &u) {
      return c == 0 ? u.x / fdtd2d::kWavesPerWG : edge_entry_c0(u.x) / (fdtd2d::kWavesPerWG * t.vw_e);
    };
    std::stable_sort(lists[c].begin(), lists[c].end(), [&](const int4 &u, const int4 &v) {
      const int gu = group(u), gv = group(v);
      if (gu != gv) return gu < gv;
      if (u.y != v.y) return u.y < v.y;
      return u.x < v.x;
    });
    for (int4 &w : lists[c]) {
      if (c == 0) {
        w.w = needs_jz(w);
        if (w.w && plasma_full(w)) w.w |= 2;
      }
      const long long c0 = c == 0 ? t.x0 + static_cast<long long>(w.x) * t.vw : edge_entry_c0(w.x);
      const long long c1 = c == 0 ? std::min<long long>(nx, c0 + t.vw) : c0 + edge_entry_w(w.x);
      const double cells = static_cast<double>(std::max<long long>(0, c1 - c0)) * (w.z - w.y);
      t.cells[c] += cells;
      if (c == 1) t.cells_kind[fdtd2d::entry_kind(w.w)] += cells;
    }
    t.n_tiles[c] = static_cast<int>(lists[c].size());
    if (c == 0)
      for (const int4 &w : lists[c]) {
        t.n_nojz += w.w == 0;
        t.n_full += (w.w & 2) != 0;
      }
    while (lists[c].size() % fdtd2d::kWavesPerWG) lists[c].push_back(make_int4(0, 0, 0, 0));
  }
  return t;
}

// The step kernel hands list positions [x*q + min(x, r), ...) of a G
// work-group launch (q = G/8, r = G%8) to XCD x in launch order. Work-groups
// with a Jz update take about 1.6 times as long as those without; moving
// them to the front of an XCD's range (keeping the order within each kind)
// leaves the short ones for the end of the launch, where they fill the
// compute units that finish early. Reorders work-groups [g0, g1).
static void jz_first_range(std::vector<int4> &list, int g0, int g1) {
  const int W = fdtd2d::kWavesPerWG;
  std::vector<int4> jz, rest;
  for (int g = g0; g < g1; ++g) {
    bool any = false;
    for (int w = 0; w < W; ++w) {
      const int4 &e = list[static_cast<size_t>(g) * W + w];
      any = any || (e.y < e.z && (e.w & 1));
    }
    std::vector<int4> &d = any ? jz : rest;
    d.insert(d.end(), list.begin() + static_cast<size_t>(g) * W, list.begin() + static_cast<size_t>(g + 1) * W);
  }
  std::copy(jz.begin(), jz.end(), list.begin() + static_cast<size_t>(g0) * W);
  std::copy(rest.begin(), rest.end(), list.begin() + static_cast<size_t>(g0) * W + jz.size());
}

static void order_jz_first(std::vector<int4> &list) {
  const int G = static_cast<int>(list.size()) / fdtd2d::kWavesPerWG, q = G >> 3, r = G & 7;
  for (int x = 0; x < 8; ++x) {
    const int g0 = x * q + std::min(x, r);
    jz_first_range(list, g0, g0 + q + (x < r ? 1 : 0));
  }
}

// One list for the fused launch: XCD x's range of the G = Gc + Gi
// work-groups (q + (x < r) of them, q = G/8, r = G%8) starts with its share
// of the x-strip and y-strip work-groups (Gc/8, one more for x < Gc%8; never
// more than the range holds), contiguous in their column order and flagged
// kFlagCpml, and continues with the next interior work-groups in column
// order (those with a Jz update first when jz_first is set). The strip
// work-groups, the longer ones per cell, start at the launch's beginning on
// every XCD.
static std::vector<int4> fused_list(const std::vector<int4> &interior, std::vector<int4> cpml, bool jz_first) {
  const int W = fdtd2d::kWavesPerWG;
  const int Gi = static_cast<int>(interior.size()) / W, Gc = static_cast<int>(cpml.size()) / W, G = Gi + Gc;
  for (int4 &e : cpml) e.w |= fdtd2d::kFlagCpml;
  std::vector<int4> out;
  out.reserve(static_cast<size_t>(G) * W);
  const int q = G >> 3, r = G & 7;
  size_t ic = 0, ii = 0;
  for (int x = 0; x < 8; ++x) {
    const int sx = q + (x < r ? 1 : 0);
    const int cx = std::min(sx, Gc / 8 + (x < Gc % 8 ? 1 : 0));
    const size_t nc = static_cast<size_t>(cx) * W, ni = static_cast<size_t>(sx - cx) * W;
    out.insert(out.end(), cpml.begin() + ic, cpml.begin() + ic + nc);
    const int g_int = static_cast<int>(out.size()) / W;
    out.insert(out.end(), interior.begin() + ii, interior.begin() + ii + ni);
    if (jz_first) jz_first_range(out, g_int, g_int + sx - cx);
    ic += nc;
    ii += ni;
  }
  return out;
}

// The fused launch is compiled for one and two columns per lane.
static bool fused_launch(const fdtd2d_plan *pl) { return pl->fused && pl->C <= 2; }

// Builds the tile lists of the plan's launch shape and puts them where the
// step kernels read them. Nothing of the plan changes unless this succeeds.
int set_tiling(fdtd2d_plan *pl) {
  const int T = pl->T > 0 ? pl->T : kDefaultT;
  TileMasks m;
  m.srow_pref = pl->srow_pref.data();
  m.scol_pref = pl->scol_pref.data();
  m.lrow_pref = pl->lrow_pref.data();
  m.pcol_pref = pl->pcol_pref.data();
  m.ne_lo = pl->ne_lo.data();
  m.ne_hi = pl->ne_hi.data();
  m.jz_lo = pl->jz_lo.data();
  if (!pl->ne_run_off.empty()) {
    m.ne_run_off = pl->ne_run_off.data();
    m.ne_runs = pl->ne_runs.data();
  }
  m.jz_hi = pl->jz_hi.data();
  m.ev = pl->ev_cells.data();
  m.nev = pl->ev_cells.size();
  Tiling t = make_tiling(pl->L.nx, pl->L.ny, pl->a.P, T, pl->C, pl->band, pl->band_cpml, &m);
  // the geometry and counts the kernels and the info calls read
  auto commit = [&](fdtd2d::StepArgs &a) {
    a.halo = t.halo;
    a.vw = t.vw;
    a.halo_e = t.halo_e;
    a.vw_e = t.vw_e;
    a.x0 = t.x0;
    a.trash_slots = std::min(kTrashSlots, kTrashRows * pl->L.pitch / (64 * pl->C));
    a.ntx = t.ntx;
    a.H = t.H;
    a.nty = t.nty;
    pl->hc = t.hc;
    for (int c = 0; c < 2; ++c) {
      pl->cells_class[c] = t.cells[c];
      a.n_tiles[c] = t.n_tiles[c];
    }
    for (int k = 0; k < 5; ++k) pl->cells_kind[k] = t.cells_kind[k];
    pl->n_nojz = t.n_nojz;
    pl->n_full = t.n_full;
  };
  fdtd2d::StepArgs &a = pl->a;
  if (pl->pooled) {  // the pool concatenates the host lists of its shots
    commit(a);
    a.n_wg[0] = static_cast<int>(t.lists[0].size()) / fdtd2d::kWavesPerWG;
    a.n_wg[1] = static_cast<int>(t.lists[1].size()) / fdtd2d::kWavesPerWG;
    a.tile_list[0] = a.tile_list[1] = nullptr;
    pl->tiling = std::move(t);
    return FDTD2D_OK;
  }
  std::vector<int4> lists[2];
  if (fused_launch(pl)) {
    // the entries of the kinds the fused kernel carries join the interior
    // ones (event entries only with no captures); the rest keep the
    // CPML-class kernel
    const bool captures = pl->a.nprobe > 0 || pl->line_rows > 0;
    std::vector<int4> xy;
    for (const int4 &e : t.lists[1]) {
      if (e.y >= e.z) continue;  // padding
      const bool in = fdtd2d::fusable_kind(fdtd2d::entry_kind(e.w), pl->a.medium_kind, captures);
      (in ? xy : lists[1]).push_back(e);
    }
    for (std::vector<int4> *v : {&xy, &lists[1]})
      while (v->size() % fdtd2d::kWavesPerWG) v->push_back(make_int4(0, 0, 0, 0));
    lists[0] = fused_list(t.lists[0], std::move(xy), pl->jz_first != 0);
  } else {
    lists[0] = std::move(t.lists[0]);
    lists[1] = std::move(t.lists[1]);
    if (pl->jz_first) order_jz_first(lists[0]);
  }
  const size_t n = lists[0].size() + lists[1].size();
  int4 *buf = pl->wg_buf;
  if (pl->static_lists) {
    // the library's static array: no allocation, a size limit
    if (n > kStaticListEntries)
      return set_error(FDTD2D_EINVAL, "the tile lists of this launch shape take " + std::to_string(n * sizeof(int4)) +
                                          " bytes, beyond the " + std::to_string(kStaticListEntries * sizeof(int4)) +
                                          "-byte static list array of fdtd2d_plan_create_in");
  } else {
    buf = nullptr;
    HIP_TRY(hipMalloc(reinterpret_cast<void **>(&buf), std::max<size_t>(n, 1) * sizeof(int4)),
            "hipMalloc tile lists");
  }
  // the launches queued so far read the current lists
  if (pl->wg_buf) {
    hipError_t e = hipStreamSynchronize(pl->stream);
    if (e != hipSuccess) {
      if (!pl->static_lists) hipFree(buf);
      return hip_fail(e, "hipStreamSynchronize");
    }
  }
  hipError_t e = hipMemcpy(buf, lists[0].data(), lists[0].size() * sizeof(int4), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(buf + lists[0].size(), lists[1].data(), lists[1].size() * sizeof(int4), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (!pl->static_lists) hipFree(buf);
    return hip_fail(e, "hipMemcpy tile lists");
  }
  if (!pl->static_lists && pl->wg_buf) hipFree(pl->wg_buf);
  pl->wg_buf = buf;
  commit(a);
  a.tile_list[0] = buf;
  a.tile_list[1] = buf + lists[0].size();
  a.n_wg[0] = static_cast<int>(lists[0].size()) / fdtd2d::kWavesPerWG;
  a.n_wg[1] = static_cast<int>(lists[1].size()) / fdtd2d::kWavesPerWG;
  return FDTD2D_OK;
}

int current_T(const fdtd2d_plan *pl) { return pl->T > 0 ? pl->T : kDefaultT; }

int upload(fdtd2d_plan *pl, size_t off, const void *src, size_t bytes) {
  if (bytes == 0) return FDTD2D_OK;
  HIP_TRY(hipMemcpy(pl->base + off, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
  return FDTD2D_OK;
}

// Per row of the problem: the runs of Ne > 0 (ne_runs, indexed by
// ne_run_off) and their column span (ne_lo / ne_hi: first and last plasma
// column, INT32_MAX / -1 for a row without plasma).
static void plasma_runs(const fdtd2d_problem *p, fdtd2d_plan *pl) {
  const int nx = p->nx, ny = p->ny;
  pl->ne_lo.assign(ny, INT32_MAX);
  pl->ne_hi.assign(ny, -1);
  pl->ne_run_off.assign(1, 0);
  pl->ne_runs.clear();
  std::vector<uint8_t> mat_pos;
  if (p->medium_kind == FDTD2D_MEDIUM_TABLE) {
    mat_pos.resize(p->nmaterials);
    for (int m = 0; m < p->nmaterials; ++m) mat_pos[m] = p->ne[m] > 0.f;
  }
  auto emit = [pl](int a, int b) { pl->ne_runs.push_back(make_int2(a, b)); };
  for (int j = 0; j < ny; ++j) {
    const size_t row = static_cast<size_t>(j) * nx;
    const size_t first = pl->ne_runs.size();
    if (p->medium_kind == FDTD2D_MEDIUM_TABLE) {
      const int32_t *mat = p->material + row;
      const uint8_t *pos = mat_pos.data();
      scan_runs(nx, [=](int i) { return static_cast<int>(pos[mat[i]]); }, emit);
    } else {
      const float *ne = p->density + row;
      scan_runs(nx, [=](int i) { return static_cast<int>(ne[i] > 0.f); }, emit);
    }
    if (pl->ne_runs.size() > first) {
      pl->ne_lo[j] = pl->ne_runs[first].x;
      pl->ne_hi[j] = pl->ne_runs.back().y - 1;
    }
    pl->ne_run_off.push_back(static_cast<int>(pl->ne_runs.size()));
  }
}

int drain_timing(fdtd2d_plan *pl) {
  for (int c = 0; c < 2; ++c) {
    for (auto &pr : pl->pending[c]) {
      float ms = 0;
      HIP_TRY(hipEventSynchronize(pr.second), "hipEventSynchronize");
      HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second), "hipEventElapsedTime");
      pl->sum_ms[c] += ms;
      pl->count[c] += 1;
      pl->spare.push_back(pr.first);
      pl->spare.push_back(pr.second);
    }
    pl->pending[c].clear();
  }
  return FDTD2D_OK;
}

hipEvent_t get_event(fdtd2d_plan *pl) {
  if (!pl->spare.empty()) {
    hipEvent_t e = pl->spare.back();
    pl->spare.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

}  // namespace detail
}  // namespace fdtd2d

extern "C" {

int fdtd2d_device_count(int *count) {
  if (!count) return set_error(FDTD2D_EINVAL, "null count");
  *count = 0;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e == hipErrorNoDevice) return FDTD2D_OK;
  HIP_TRY(e, "hipGetDeviceCount");
  *count = n;
  return FDTD2D_OK;
}

int fdtd2d_tile_lists(int nx, int ny, int npml, int tblock, int cols_per_lane, int band, int band_cpml,
                      const int *ev_i, const int *ev_j, int nev, int *entries, size_t capacity, size_t *count,
                      int *geometry) {
  if (!count) return set_error(FDTD2D_EINVAL, "null count");
  if (nx < 3 || ny < 3 || npml < 0 || 2 * npml > nx - 2 || 2 * npml > ny - 2)
    return set_error(FDTD2D_EINVAL, "grid or CPML size out of range");
  const int T = tblock > 0 ? tblock : kDefaultT;
  if (T > fdtd2d::kMaxT) return set_error(FDTD2D_EINVAL, "tblock must be 1..8");
  if (cols_per_lane != 1 && cols_per_lane != 2 && cols_per_lane != 4)
    return set_error(FDTD2D_EINVAL, "cols_per_lane must be 1, 2 or 4");
  if (band < 0 || band_cpml < 0) return set_error(FDTD2D_EINVAL, "negative band");
  if (nev < 0 || (nev > 0 && (!ev_i || !ev_j))) return set_error(FDTD2D_EINVAL, "bad event cells");
  std::vector<int2> ev;
  for (int e = 0; e < nev; ++e) ev.push_back(make_int2(ev_i[e], ev_j[e]));
  std::sort(ev.begin(), ev.end(), [](const int2 &u, const int2 &v) { return u.y != v.y ? u.y < v.y : u.x < v.x; });
  TileMasks m;
  m.ev = ev.data();
  m.nev = ev.size();
  const Tiling t = make_tiling(nx, ny, npml, T, cols_per_lane, band, band_cpml, &m);
  *count = static_cast<size_t>(t.n_tiles[0]) + t.n_tiles[1];
  if (geometry) {
    geometry[0] = t.vw;
    geometry[1] = t.halo;
    geometry[2] = t.H;
    geometry[3] = t.hc;
  }
  if (!entries) return FDTD2D_OK;
  if (capacity < *count) return set_error(FDTD2D_EINVAL, "entries too small");
  size_t k = 0;
  for (int c = 0; c < 2; ++c)
    for (int q = 0; q < t.n_tiles[c]; ++q, ++k) {
      const int4 &w = t.lists[c][q];
      const int c0 = c == 0 ? t.x0 + w.x * t.vw : edge_entry_c0(w.x);
      const int c1 = c == 0 ? std::min(nx, c0 + t.vw) : c0 + edge_entry_w(w.x);
      int *e = entries + 5 * k;
      e[0] = c0;
      e[1] = c1;
      e[2] = w.y;
      e[3] = w.z;
      e[4] = c == 0 ? 0 : 1 + fdtd2d::entry_kind(w.w);
    }
  return FDTD2D_OK;
}

int fdtd2d_plan_footprint(const fdtd2d_problem *p, uint64_t *bytes) {
  if (!bytes) return set_error(FDTD2D_EINVAL, "null bytes");
  if (!p) return set_error(FDTD2D_EINVAL, "null problem");
  if (p->nx < 4 || p->ny < 4 || p->nsteps < 0 || p->npml < 0)
    return set_error(FDTD2D_EINVAL, "bad sizes");
  size_t nev = static_cast<size_t>(p->nsrc) + p->ndet + p->ncomp;
  *bytes = make_layout(p, nev).total;
  return FDTD2D_OK;
}

int fdtd2d_plan_create(const fdtd2d_problem *p, int device, fdtd2d_plan **out) {
  const int C = p ? default_cols_per_lane(p->nx, p->ny) : 1;
  return create_plan(p, device, false, nullptr, 0, C, 0, 0, nullptr, out);
}

int fdtd2d_plan_create_in(const fdtd2d_problem *p, int device, void *mem, size_t bytes, void *stream,
                          fdtd2d_plan **out) {
  if (!out) return set_error(FDTD2D_EINVAL, "null out");
  *out = nullptr;
  if (!p) return set_error(FDTD2D_EINVAL, "null problem");
  if (!mem) return set_error(FDTD2D_EINVAL, "null device memory");
  if (reinterpret_cast<uintptr_t>(mem) % kAlign)
    return set_error(FDTD2D_EINVAL, "the plan's memory must be 256-byte aligned");
  int rc = fdtd2d::validate_problem(p);
  if (rc) return rc;
  const size_t need = plan_bytes(p);
  if (bytes < need)
    return set_error(FDTD2D_EINVAL, "the plan needs " + std::to_string(need) + " bytes of device memory (" +
                                        std::to_string(bytes) + " given; fdtd2d_plan_footprint)");
  if (device < 0 || device >= kMaxDevices) return set_error(FDTD2D_EINVAL, "no such device");
  {
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (g_static_busy[device])
      return set_error(FDTD2D_ESTATE, "another plan made by fdtd2d_plan_create_in exists on this device (they "
                                      "share its static tile-list array)");
    g_static_busy[device] = true;
  }
  const int C = default_cols_per_lane(p->nx, p->ny);
  rc = create_plan(p, device, false, static_cast<hipStream_t>(stream), 0, C, 0, 0, static_cast<char *>(mem), out,
                   true);
  std::lock_guard<std::mutex> lk(g_static_mu);
  if (rc) g_static_busy[device] = false;  // a plan holds the array only once it exists
  else (*out)->holds_static = true;
  return rc;
}

}  // extern "C"

size_t fdtd2d::detail::plan_bytes(const fdtd2d_problem *p) {
  return make_layout(p, static_cast<size_t>(p->nsrc) + p->ndet + p->ncomp).total;
}

int fdtd2d::detail::create_plan(const fdtd2d_problem *p, int device, bool pooled, hipStream_t stream, int T, int C,
                                int band, int band_cpml, char *base, fdtd2d_plan **out, bool caller_mem) {
  if (!out) return set_error(FDTD2D_EINVAL, "null out");
  *out = nullptr;
  int rc = fdtd2d::validate_problem(p);
  if (rc) return rc;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev), "hipGetDeviceCount");
  if (device < 0 || device >= ndev) return set_error(FDTD2D_EINVAL, "no such device");
  HIP_TRY(hipSetDevice(device), "hipSetDevice");

  if (static_cast<long long>(align_up(p->nx, 64)) * (p->ny + kTrashRows) * 4 > 0x7fffffffLL)
    return set_error(FDTD2D_EUNSUPPORTED, "a field plane above 2 GiB (32-bit offsets)");
  std::vector<int32_t> ev_off;
  std::vector<fdtd2d::Event> events = build_events(p, ev_off);
  auto *pl = new (std::nothrow) fdtd2d_plan();
  if (!pl) return set_error(FDTD2D_ENOMEM, "out of host memory");
  pl->device = device;
  pl->pooled = pooled;
  pl->T = T;
  pl->C = C;
  // rows of prefetch: one with wider tiles (the fused launch at 8192^2: 838
  // against 829 Gcell/s with two; 16384^2 835 against 827), two otherwise
  pl->D = C >= 2 ? 1 : 2;
  pl->band = band;
  pl->band_cpml = band_cpml;
  pl->L = make_layout(p, events.size());
  pl->nsteps = p->nsteps;
  pl->ndet = p->ndet;
  pl->ncomp = p->ncomp;
  pl->nsrc = p->nsrc;
  const Layout &L = pl->L;
  hipError_t e = hipSuccess;
  if (base) {
    pl->base = base;
    pl->owns_base = false;
  } else {
    e = hipMalloc(reinterpret_cast<void **>(&pl->base), L.total);
  }
  if (e != hipSuccess) {
    delete pl;
    return set_error(FDTD2D_ENOMEM, std::string("hipMalloc of ") + std::to_string(L.total) +
                                        " bytes: " + hipGetErrorString(e));
  }
  auto fail = [&](int code) {
    fdtd2d_plan_destroy(pl);
    return code;
  };
  if (caller_mem) {
    // the caller's memory and stream; tile lists in the static array
    pl->stream = stream;
    pl->static_lists = true;
    rc = static_lists_ptr(device, &pl->wg_buf);
    if (rc) return fail(rc);
  } else if (pooled) {
    // a pool orders its shots' work itself (events on its own streams); the
    // plan's copies go to the null stream, which never waits for the pool's
    (void)stream;
    pl->stream = nullptr;
  } else {
    if (hipStreamCreateWithFlags(&pl->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&pl->aux, hipStreamNonBlocking) != hipSuccess)
      return fail(set_error(FDTD2D_EDEVICE, "hipStreamCreate"));
    if (hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming) != hipSuccess)
      return fail(set_error(FDTD2D_EDEVICE, "hipEventCreate"));
    pl->stream = pl->own_stream;
  }
  if (caller_mem) {
    // L.total is a multiple of 256 bytes (make_layout)
    hipLaunchKernelGGL(zero_words, dim3(1024), dim3(256), 0, pl->stream, reinterpret_cast<uint4 *>(pl->base),
                       L.total / 16);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(pl->stream) != hipSuccess)
      return fail(set_error(FDTD2D_EDEVICE, "zeroing the plan's memory"));
  } else if (hipMemset(pl->base, 0, L.total) != hipSuccess) {
    return fail(set_error(FDTD2D_EDEVICE, "hipMemset"));
  }

  const int nx = L.nx, ny = L.ny, pitch = L.pitch, P = L.P;
  // the tables from the material table to the source table are staged in one
  // host image and copied with one transfer (gaps between them are zero)
  std::vector<char> img(L.off_det - L.off_lut, 0);
  auto stage = [&](size_t off, const void *src, size_t bytes) {
    if (bytes) std::memcpy(img.data() + (off - L.off_lut), src, bytes);
    return FDTD2D_OK;
  };
  // medium in pitch layout: the material ids or the density bits, 4 bytes a
  // cell either way (the pad columns stay zero)
  {
    const void *cells = p->medium_kind == FDTD2D_MEDIUM_TABLE ? static_cast<const void *>(p->material)
                                                              : static_cast<const void *>(p->density);
    if (pitch == nx) {
      if ((rc = upload(pl, L.off_med, cells, static_cast<size_t>(ny) * nx * 4))) return fail(rc);
    } else {
      std::vector<uint32_t> med(static_cast<size_t>(ny) * pitch, 0u);
      const char *src = static_cast<const char *>(cells);
      for (int j = 0; j < ny; ++j)
        std::memcpy(&med[static_cast<size_t>(j) * pitch], src + static_cast<size_t>(j) * nx * 4,
                    static_cast<size_t>(nx) * 4);
      if ((rc = upload(pl, L.off_med, med.data(), med.size() * 4))) return fail(rc);
    }
  }
  if (p->medium_kind == FDTD2D_MEDIUM_TABLE) {
    std::vector<float> lut(static_cast<size_t>(p->nmaterials) * 4, 0.f);
    for (int m = 0; m < p->nmaterials; ++m) {
      lut[4 * m + 0] = p->ca[m];
      lut[4 * m + 1] = p->cb[m];
      lut[4 * m + 2] = p->ne[m];
    }
    if ((rc = stage(L.off_lut, lut.data(), lut.size() * 4))) return fail(rc);
  }
  // per-column tables: den_ex, den_hx, hyx_b, hyx_c, ezx_b, ezx_c, hyx_slot, ezx_slot
  {
    std::vector<float> col[6];
    for (auto &v : col) v.assign(pitch, 0.f);
    std::vector<int32_t> hyx_s(pitch, -1), ezx_s(pitch, -1);
    for (int i = 0; i < nx - 1; ++i) {
      col[0][i] = p->den_ex[i];
      col[1][i] = p->den_hx[i];
    }
    if (P > 0) {
      for (int r = 0; r <= P - 2; ++r) {  // Hy x strips: i = r and nx-2-r
        hyx_s[r] = r;
        col[2][r] = p->bh_x[0][r];
        col[3][r] = p->ch_x[0][r];
        const int i2 = nx - 2 - r;
        hyx_s[i2] = P + r;
        col[2][i2] = p->bh_x[1][r];
        col[3][i2] = p->ch_x[1][r];
      }
      for (int r = 1; r <= P - 1; ++r) {  // Ez x strips: i = r and nx-1-r
        ezx_s[r] = r;
        col[4][r] = p->be_x[0][r];
        col[5][r] = p->ce_x[0][r];
        const int i2 = nx - 1 - r;
        ezx_s[i2] = P + r;
        col[4][i2] = p->be_x[1][r];
        col[5][i2] = p->ce_x[1][r];
      }
    }
    for (int f = 0; f < 6; ++f)
      if ((rc = stage(L.off_col[f], col[f].data(), pitch * 4))) return fail(rc);
    if ((rc = stage(L.off_col[6], hyx_s.data(), pitch * 4))) return fail(rc);
    if ((rc = stage(L.off_col[7], ezx_s.data(), pitch * 4))) return fail(rc);
  }
  // per-row tables: den_ey, den_hy, hxy_b, hxy_c, ezy_b, ezy_c, hxy_slot, ezy_slot
  {
    std::vector<float> row[6];
    for (auto &v : row) v.assign(ny, 0.f);
    std::vector<int32_t> hxy_s(ny, -1), ezy_s(ny, -1);
    for (int j = 0; j < ny - 1; ++j) {
      row[0][j] = p->den_ey[j];
      row[1][j] = p->den_hy[j];
    }
    if (P > 0) {
      for (int r = 0; r <= P - 2; ++r) {
        hxy_s[r] = r;
        row[2][r] = p->bh_y[0][r];
        row[3][r] = p->ch_y[0][r];
        const int j2 = ny - 2 - r;
        hxy_s[j2] = P + r;
        row[2][j2] = p->bh_y[1][r];
        row[3][j2] = p->ch_y[1][r];
      }
      for (int r = 1; r <= P - 1; ++r) {
        ezy_s[r] = r;
        row[4][r] = p->be_y[0][r];
        row[5][r] = p->ce_y[0][r];
        const int j2 = ny - 1 - r;
        ezy_s[j2] = P + r;
        row[4][j2] = p->be_y[1][r];
        row[5][j2] = p->ce_y[1][r];
      }
    }
    for (int f = 0; f < 6; ++f)
      if ((rc = stage(L.off_row[f], row[f].data(), ny * 4))) return fail(rc);
    if ((rc = stage(L.off_row[6], hxy_s.data(), ny * 4))) return fail(rc);
    if ((rc = stage(L.off_row[7], ezy_s.data(), ny * 4))) return fail(rc);
    // the per-row record the CPML-class kernel reads at each level
    // (kMaxT records on each side for the rows the levels reach outside the
    // grid: no events, no strips)
    const int pad = fdtd2d::kMaxT;
    std::vector<int32_t> ri(static_cast<size_t>(ny + 2 * pad) * fdtd2d::kRowRec, 0);
    for (int j = -pad; j < ny + pad; ++j) {
      int32_t *q = &ri[static_cast<size_t>(j + pad) * fdtd2d::kRowRec];
      q[2] = -1;
      q[4] = q[5] = -1;
    }
    auto bits = [](float v) {
      int32_t b;
      std::memcpy(&b, &v, 4);
      return b;
    };
    for (int j = 0; j < ny; ++j) {
      const int jc = std::min(j, ny - 2);
      int32_t *q = &ri[static_cast<size_t>(j + pad) * fdtd2d::kRowRec];
      q[0] = bits(p->den_hy[jc]);
      q[1] = bits(p->den_ey[jc]);
      q[2] = ev_off[j] < ev_off[j + 1] ? ev_off[j] : -1;
      q[3] = ev_off[j + 1];
      const bool hrow = j <= ny - 2;  // psi updates act on rows 0..ny-2 only
      q[4] = hrow ? hxy_s[j] : -1;
      q[5] = hrow ? ezy_s[j] : -1;
      q[6] = bits(hrow ? row[2][j] : 0.f);
      q[7] = bits(hrow ? row[3][j] : 0.f);
      q[8] = bits(hrow ? row[4][j] : 0.f);
      q[9] = bits(hrow ? row[5][j] : 0.f);
    }
    if ((rc = stage(L.off_rowinfo, ri.data(), ri.size() * 4))) return fail(rc);
  }
  // uniform interior den values (bitwise) and the rows / columns that differ
  float u_hx, u_ex, u_hy, u_ey;
  {
    const int im = (nx - 1) / 2, jm = (ny - 1) / 2;
    u_hx = p->den_hx[im];
    u_ex = p->den_ex[im];
    u_hy = p->den_hy[jm];
    u_ey = p->den_ey[jm];
    auto same = [](float x, float y) { return std::memcmp(&x, &y, 4) == 0; };
    // RC probe columns and the processing line's row are special too: only
    // general CPML-class tiles capture them
    std::vector<char> probe(nx, 0);
    for (int e = 0; e < p->nrc; ++e) probe[p->rc_i[e]] = probe[p->rc_imax[e]] = 1;
    pl->scol_pref.assign(nx + 1, 0);
    pl->pcol_pref.assign(nx + 1, 0);
    for (int i = 0; i < nx; ++i) {
      const bool sp = i >= nx - 1 || !same(p->den_hx[i], u_hx) || !same(p->den_ex[i], u_ex);
      pl->scol_pref[i + 1] = pl->scol_pref[i] + sp;
      pl->pcol_pref[i + 1] = pl->pcol_pref[i] + (probe[i] != 0);
    }
    pl->srow_pref.assign(ny + 1, 0);
    pl->lrow_pref.assign(ny + 1, 0);
    for (int j = 0; j < ny; ++j) {
      const bool sp = j >= ny - 1 || !same(p->den_hy[j], u_hy) || !same(p->den_ey[j], u_ey);
      pl->srow_pref[j + 1] = pl->srow_pref[j] + sp;
      pl->lrow_pref[j + 1] = pl->lrow_pref[j] + (L.line_rows > 0 && j == p->line_j);
    }
  }
  {
    for (int e = 0; e < p->nsrc; ++e) pl->ev_cells.push_back(make_int2(p->src_i[e], p->src_j[e]));
    for (int e = 0; e < p->ndet; ++e) pl->ev_cells.push_back(make_int2(p->det_i[e], p->det_j[e]));
    for (int e = 0; e < p->ncomp; ++e) pl->ev_cells.push_back(make_int2(p->comp_i[e], p->comp_j[e]));
    std::sort(pl->ev_cells.begin(), pl->ev_cells.end(),
              [](const int2 &u, const int2 &v) { return u.y != v.y ? u.y < v.y : u.x < v.x; });
  }
  // column spans of plasma per row; Jz starts at +0 everywhere
  plasma_runs(p, pl);
  pl->jz_lo.assign(ny, INT32_MAX);
  pl->jz_hi.assign(ny, -1);
  if ((rc = stage(L.off_evoff, ev_off.data(), ev_off.size() * 4))) return fail(rc);
  if ((rc = stage(L.off_events, events.data(), events.size() * sizeof(fdtd2d::Event)))) return fail(rc);
  // RC tracker: probe index per column, per-line probe pair, edge levels
  {
    std::vector<int32_t> pcol(pitch, -1);
    std::vector<int> cols;
    for (int e = 0; e < p->nrc; ++e) {
      cols.push_back(p->rc_i[e]);
      cols.push_back(p->rc_imax[e]);
    }
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    for (size_t q = 0; q < cols.size(); ++q) pcol[cols[q]] = static_cast<int32_t>(q);
    std::vector<int32_t> map(2 * static_cast<size_t>(p->nrc));
    for (int e = 0; e < p->nrc; ++e) {
      map[2 * e] = pcol[p->rc_imax[e]];
      map[2 * e + 1] = pcol[p->rc_i[e]];
    }
    std::vector<float> edge(p->nrc, p->rc_edge0);
    if ((rc = upload(pl, L.off_probe, pcol.data(), pcol.size() * 4))) return fail(rc);
    if ((rc = upload(pl, L.off_rcmap, map.data(), map.size() * 4))) return fail(rc);
    if ((rc = upload(pl, L.off_rcedge, edge.data(), edge.size() * 4))) return fail(rc);
  }
  if ((rc = stage(L.off_src, p->src_table, static_cast<size_t>(p->nsteps + 1) * p->nsrc * 4)))
    return fail(rc);
  if ((rc = upload(pl, L.off_lut, img.data(), img.size()))) return fail(rc);

  fdtd2d::StepArgs &a = pl->a;
  a.nx = nx;
  a.ny = ny;
  a.pitch = pitch;
  a.trash_rows = kTrashRows;
  a.P = P;
  a.xs = L.xs;
  a.nsteps = p->nsteps;
  a.medium_kind = p->medium_kind;
  a.nmat = p->medium_kind == FDTD2D_MEDIUM_TABLE ? p->nmaterials : 0;
  a.pec = p->medium_kind == FDTD2D_MEDIUM_TABLE ? p->pec_material : -1;
  a.db = p->db;
  a.ca0 = p->ca0;
  a.cb0 = p->cb0;
  a.cj = p->cj;
  a.jz_decay = p->jz_decay;
  a.dxf = p->dx_f;
  a.dyf = p->dy_f;
  a.rdx = 1.0 / static_cast<double>(p->dx_f);
  a.rdy = 1.0 / static_cast<double>(p->dy_f);
  a.jz_dt = p->jz_dt;
  a.uden_hx = u_hx;
  a.uden_ex = u_ex;
  a.uden_hy = u_hy;
  a.uden_ey = u_ey;
  for (int s = 0; s < 2; ++s) {
    a.ez[s] = reinterpret_cast<float *>(pl->base + L.off_fields[s][0]);
    a.hx[s] = reinterpret_cast<float *>(pl->base + L.off_fields[s][1]);
    a.hy[s] = reinterpret_cast<float *>(pl->base + L.off_fields[s][2]);
    a.jz[s] = reinterpret_cast<float *>(pl->base + L.off_fields[s][3]);
    a.psi_hxy[s] = reinterpret_cast<float *>(pl->base + L.off_psi_y[s][0]);
    a.psi_ezy[s] = reinterpret_cast<float *>(pl->base + L.off_psi_y[s][1]);
    a.psi_hyx[s] = reinterpret_cast<float *>(pl->base + L.off_psi_x[s][0]);
    a.psi_ezx[s] = reinterpret_cast<float *>(pl->base + L.off_psi_x[s][1]);
  }
  a.med_id = reinterpret_cast<const int32_t *>(pl->base + L.off_med);
  a.med_ne = reinterpret_cast<const float *>(pl->base + L.off_med);
  a.lut = reinterpret_cast<const float4 *>(pl->base + L.off_lut);
  a.den_ex = reinterpret_cast<const float *>(pl->base + L.off_col[0]);
  a.den_hx = reinterpret_cast<const float *>(pl->base + L.off_col[1]);
  a.hyx_b = reinterpret_cast<const float *>(pl->base + L.off_col[2]);
  a.hyx_c = reinterpret_cast<const float *>(pl->base + L.off_col[3]);
  a.ezx_b = reinterpret_cast<const float *>(pl->base + L.off_col[4]);
  a.ezx_c = reinterpret_cast<const float *>(pl->base + L.off_col[5]);
  a.hyx_slot = reinterpret_cast<const int32_t *>(pl->base + L.off_col[6]);
  a.ezx_slot = reinterpret_cast<const int32_t *>(pl->base + L.off_col[7]);
  a.den_ey = reinterpret_cast<const float *>(pl->base + L.off_row[0]);
  a.den_hy = reinterpret_cast<const float *>(pl->base + L.off_row[1]);
  a.hxy_b = reinterpret_cast<const float *>(pl->base + L.off_row[2]);
  a.hxy_c = reinterpret_cast<const float *>(pl->base + L.off_row[3]);
  a.ezy_b = reinterpret_cast<const float *>(pl->base + L.off_row[4]);
  a.ezy_c = reinterpret_cast<const float *>(pl->base + L.off_row[5]);
  a.hxy_slot = reinterpret_cast<const int32_t *>(pl->base + L.off_row[6]);
  a.ezy_slot = reinterpret_cast<const int32_t *>(pl->base + L.off_row[7]);
  a.ev_off = reinterpret_cast<const int32_t *>(pl->base + L.off_evoff);
  a.rowinfo = reinterpret_cast<const int *>(pl->base + L.off_rowinfo) + fdtd2d::kMaxT * fdtd2d::kRowRec;
  a.events = reinterpret_cast<const fdtd2d::Event *>(pl->base + L.off_events);
  a.src_table = reinterpret_cast<const float *>(pl->base + L.off_src);
  a.nsrc = p->nsrc;
  a.det = reinterpret_cast<float *>(pl->base + L.off_det);
  a.ndet = p->ndet;
  a.comp = reinterpret_cast<float *>(pl->base + L.off_comp);
  a.ncomp = p->ncomp;
  a.comp_plane = static_cast<long long>(p->nsteps + 1) * p->ncomp;
  a.probe_col = reinterpret_cast<const int *>(pl->base + L.off_probe);
  a.rc_cap = reinterpret_cast<float *>(pl->base + L.off_cap);
  a.nprobe = L.nprobe;
  pl->nrc = p->nrc;
  pl->rc.cap = a.rc_cap;
  pl->rc.map = reinterpret_cast<const int *>(pl->base + L.off_rcmap);
  pl->rc.edge = reinterpret_cast<float *>(pl->base + L.off_rcedge);
  pl->rc.spd = reinterpret_cast<int *>(pl->base + L.off_spd);
  pl->rc.nrc = p->nrc;
  pl->rc.nprobe = L.nprobe;
  pl->rc.ny = ny;
  pl->rc.jcut = p->rc_jcut;
  pl->rc.level = p->rc_level;
  pl->line_rows = L.line_rows;
  pl->line_size = L.line_rows > 0 ? p->line_size : 0;
  a.line = reinterpret_cast<float *>(pl->base + L.off_line);
  a.line_j = L.line_rows > 0 ? p->line_j : -1;
  a.line_i0 = p->line_i0;
  a.line_i1 = p->line_i0 + pl->line_size;
  a.line_begin = p->line_begin;
  a.line_end = std::min(p->line_end, p->nsteps);
  if ((rc = set_tiling(pl))) return fail(rc);
  *out = pl;
  return FDTD2D_OK;
}

extern "C" {

int fdtd2d_plan_set_stream(fdtd2d_plan *pl, void *stream) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  pl->stream = stream ? static_cast<hipStream_t>(stream) : pl->own_stream;
  return FDTD2D_OK;
}

int fdtd2d_plan_set_tblock(fdtd2d_plan *pl, int T) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  if (T < 0 || T > fdtd2d::kMaxT) return set_error(FDTD2D_EINVAL, "steps per launch must be 0..8");
  const int old = pl->T;
  pl->T = T;
  const int rc = set_tiling(pl);
  if (rc) pl->T = old;
  return rc;
}

int fdtd2d_plan_set_option(fdtd2d_plan *pl, const char *name, double value) {
  if (!pl || !name) return set_error(FDTD2D_EINVAL, "null argument");
  std::string n(name);
  const int v = static_cast<int>(value);
  // the shape before the change, restored if the new one cannot be set up
  const int oT = pl->T, oband = pl->band, oband_cpml = pl->band_cpml, oC = pl->C, oD = pl->D,
            ojz = pl->jz_first, ofused = pl->fused;
  auto retile = [&]() {
    const int rc = set_tiling(pl);
    if (rc) {
      pl->T = oT;
      pl->band = oband;
      pl->band_cpml = oband_cpml;
      pl->C = oC;
      pl->D = oD;
      pl->jz_first = ojz;
      pl->fused = ofused;
    }
    return rc;
  };
  if (n == "tblock") {
    if (v < 0 || v > fdtd2d::kMaxT) return set_error(FDTD2D_EINVAL, "tblock out of range");
    pl->T = v;
  } else if (n == "band") {
    if (v < 0 || v > pl->L.ny) return set_error(FDTD2D_EINVAL, "band out of range");
    pl->band = v;
  } else if (n == "band_cpml") {
    if (v < 0 || v > pl->L.ny) return set_error(FDTD2D_EINVAL, "band_cpml out of range");
    pl->band_cpml = v;
  } else if (n == "cols_per_lane") {
    if (v != 1 && v != 2 && v != 4) return set_error(FDTD2D_EINVAL, "cols_per_lane must be 1, 2 or 4");
    pl->C = v;
    if (v == 4) pl->D = 1;
    if (v == 2 && pl->D > 2) pl->D = 2;
  } else if (n == "jz_first") {
    if (v != 0 && v != 1) return set_error(FDTD2D_EINVAL, "jz_first must be 0 or 1");
    pl->jz_first = v;
  } else if (n == "fused") {
    if (v != 0 && v != 1) return set_error(FDTD2D_EINVAL, "fused must be 0 or 1");
    pl->fused = v;
  } else if (n == "prefetch") {
    const int maxd = pl->C == 1 ? 4 : pl->C == 2 ? 2 : 1;
    if (v != 1 && v != 2 && v != 4) return set_error(FDTD2D_EINVAL, "prefetch must be 1, 2 or 4");
    if (v > maxd) return set_error(FDTD2D_EINVAL, "prefetch depth too large for cols_per_lane");
    pl->D = v;
  } else {
    return set_error(FDTD2D_EINVAL, "unknown option " + n);
  }
  return retile();
}

int fdtd2d_plan_get_info(const fdtd2d_plan *pl, const char *name, double *value) {
  if (!pl || !name || !value) return set_error(FDTD2D_EINVAL, "null argument");
  std::string n(name);
  if (n == "pitch") *value = pl->L.pitch;
  else if (n == "tblock") *value = current_T(pl);
  else if (n == "band") *value = pl->a.H;
  else if (n == "band_cpml") *value = pl->hc;
  else if (n == "ntx") *value = pl->a.ntx;
  else if (n == "nty") *value = pl->a.nty;
  else if (n == "workgroups") *value = pl->a.n_wg[0] + pl->a.n_wg[1];
  else if (n == "workgroups_cpml") *value = pl->a.n_wg[1];
  else if (n == "tiles_interior") *value = pl->a.n_tiles[0];
  else if (n == "tiles_cpml") *value = pl->a.n_tiles[1];
  else if (n == "tiles_nojz") *value = pl->n_nojz;
  else if (n == "tiles_plasma_full") *value = pl->n_full;
  else if (n == "cells_interior") *value = pl->cells_class[0];
  else if (n == "cells_cpml") *value = pl->cells_class[1];
  else if (n == "cells_cpml_general") *value = pl->cells_kind[fdtd2d::kKindGeneral];
  else if (n == "cells_xstrip") *value = pl->cells_kind[fdtd2d::kKindX];
  else if (n == "cells_ystrip") *value = pl->cells_kind[fdtd2d::kKindY];
  else if (n == "cells_corner") *value = pl->cells_kind[fdtd2d::kKindXY];
  else if (n == "cells_events") *value = pl->cells_kind[fdtd2d::kKindE];
  else if (n == "tile_cols") *value = pl->a.vw;
  else if (n == "bytes") *value = static_cast<double>(pl->L.total);
  else if (n == "steps") *value = pl->n;
  else if (n == "cols_per_lane") *value = pl->C;
  else if (n == "prefetch") *value = pl->D;
  else if (n == "jz_first") *value = pl->jz_first;
  else if (n == "fused") *value = fused_launch(pl) ? 1 : 0;
  else if (n == "cus") {
    int ncu = 0;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, pl->device), "hipDeviceGetAttribute");
    *value = ncu;
  } else return set_error(FDTD2D_EINVAL, "unknown info name " + n);
  return FDTD2D_OK;
}

int fdtd2d_run(fdtd2d_plan *pl, int nsteps) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  if (pl->pooled) return set_error(FDTD2D_ESTATE, "a pool's shot is stepped by the pool");
  if (nsteps < 0 || pl->n + nsteps > pl->nsteps)
    return set_error(FDTD2D_EINVAL, "run beyond the problem's nsteps (" + std::to_string(pl->nsteps) + ")");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  const int Tmax = current_T(pl);
  while (nsteps > 0) {
    const int t = std::min(Tmax, nsteps);
    hipEvent_t e0 = nullptr, e1 = nullptr, c0 = nullptr, c1 = nullptr;
    if (pl->timing) {
      e0 = get_event(pl);
      e1 = get_event(pl);
      c0 = get_event(pl);
      c1 = get_event(pl);
      if (!e0 || !e1 || !c0 || !c1) return set_error(FDTD2D_EDEVICE, "hipEventCreate");
    }
    // CPML work-groups on the aux stream, interior ones on the main stream;
    // both read buffer set src and write the other, so they may overlap. A
    // plan with no aux stream (fdtd2d_plan_create_in) runs the two kernels
    // one after the other on its stream.
    if (pl->a.n_wg[1] > 0 && !pl->aux) {
      if (pl->timing) HIP_TRY(hipEventRecord(c0, pl->stream), "hipEventRecord");
      hipError_t e = fdtd2d::launch_step(pl->a, fused_launch(pl) ? 2 : 1, t, pl->C, 1, pl->src, pl->n, pl->stream);
      if (e != hipSuccess) return hip_fail(e, "CPML step kernel launch");
      if (pl->timing) {
        HIP_TRY(hipEventRecord(c1, pl->stream), "hipEventRecord");
        pl->pending[1].push_back({c0, c1});
      }
    } else if (pl->a.n_wg[1] > 0) {
      HIP_TRY(hipEventRecord(pl->ev_fork, pl->stream), "hipEventRecord fork");
      HIP_TRY(hipStreamWaitEvent(pl->aux, pl->ev_fork, 0), "hipStreamWaitEvent fork");
      if (pl->timing) HIP_TRY(hipEventRecord(c0, pl->aux), "hipEventRecord");
      hipError_t e = fdtd2d::launch_step(pl->a, fused_launch(pl) ? 2 : 1, t, pl->C, 1, pl->src, pl->n, pl->aux);
      if (e != hipSuccess) return hip_fail(e, "CPML step kernel launch");
      if (pl->timing) {
        HIP_TRY(hipEventRecord(c1, pl->aux), "hipEventRecord");
        pl->pending[1].push_back({c0, c1});
      }
      HIP_TRY(hipEventRecord(pl->ev_join, pl->aux), "hipEventRecord join");
    } else if (pl->timing) {
      pl->spare.push_back(c0);
      pl->spare.push_back(c1);
    }
    // timing events bracket each kernel alone on its own stream
    if (pl->timing) HIP_TRY(hipEventRecord(e0, pl->stream), "hipEventRecord");
    hipError_t e = fused_launch(pl) ? fdtd2d::launch_step_fused(pl->a, t, pl->C, pl->D, pl->src, pl->n, pl->stream)
                             : fdtd2d::launch_step(pl->a, 0, t, pl->C, pl->D, pl->src, pl->n, pl->stream);
    if (e != hipSuccess) return hip_fail(e, "step kernel launch");
    if (pl->timing) {
      HIP_TRY(hipEventRecord(e1, pl->stream), "hipEventRecord");
      pl->pending[0].push_back({e0, e1});
    }
    if (pl->a.n_wg[1] > 0 && pl->aux) HIP_TRY(hipStreamWaitEvent(pl->stream, pl->ev_join, 0), "hipStreamWaitEvent join");
    if (pl->nrc > 0) {
      e = fdtd2d::launch_rc(pl->rc, pl->n, t, pl->stream);
      if (e != hipSuccess) return hip_fail(e, "RC kernel launch");
    }
    pl->launches_T[t] += 1;
    pl->src = 1 - pl->src;
    pl->n += t;
    nsteps -= t;
  }
  return FDTD2D_OK;
}

int fdtd2d_sync(fdtd2d_plan *pl) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
  return drain_timing(pl);
}

int fdtd2d_steps_done(const fdtd2d_plan *pl, int *n) {
  if (!pl || !n) return set_error(FDTD2D_EINVAL, "null argument");
  *n = pl->n;
  return FDTD2D_OK;
}

int fdtd2d_read_detectors(fdtd2d_plan *pl, float *det, size_t det_count, float *comp, size_t comp_count) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
  const size_t nd = static_cast<size_t>(pl->nsteps + 1) * pl->ndet;
  const size_t nc = static_cast<size_t>(3) * (pl->nsteps + 1) * pl->ncomp;
  if (det) {
    if (det_count != nd) return set_error(FDTD2D_EINVAL, "det_count must be (nsteps+1)*ndet");
    if (nd) HIP_TRY(hipMemcpy(det, pl->a.det, nd * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H det");
  }
  if (comp) {
    if (comp_count != nc) return set_error(FDTD2D_EINVAL, "comp_count must be 3*(nsteps+1)*ncomp");
    if (nc) HIP_TRY(hipMemcpy(comp, pl->a.comp, nc * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H comp");
  }
  return FDTD2D_OK;
}

int fdtd2d_read_rc(fdtd2d_plan *pl, int32_t *spd, size_t count) {
  if (!pl || !spd) return set_error(FDTD2D_EINVAL, "null argument");
  const size_t n = static_cast<size_t>(pl->nsteps + 1) * pl->nrc;
  if (count != n) return set_error(FDTD2D_EINVAL, "count must be (nsteps+1)*nrc");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
  if (n) HIP_TRY(hipMemcpy(spd, pl->rc.spd, n * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H rc");
  return FDTD2D_OK;
}

int fdtd2d_read_line(fdtd2d_plan *pl, float *line, size_t count) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  const size_t n = pl->line_rows * pl->line_size;
  if (n && !line) return set_error(FDTD2D_EINVAL, "null line buffer");
  if (count != n) return set_error(FDTD2D_EINVAL, "count must be rows*line_size of the processing line");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
  if (n) HIP_TRY(hipMemcpy(line, pl->a.line, n * 4, hipMemcpyDeviceToHost), "hipMemcpy D2H line");
  return FDTD2D_OK;
}

static int field_shape(const fdtd2d_plan *pl, int which, int &rows, int &cols, float *const **bufs) {
  const int nx = pl->L.nx, ny = pl->L.ny;
  switch (which) {
    case FDTD2D_EZ: rows = ny; cols = nx; *bufs = pl->a.ez; return 0;
    case FDTD2D_HX: rows = ny - 1; cols = nx; *bufs = pl->a.hx; return 0;
    case FDTD2D_HY: rows = ny; cols = nx - 1; *bufs = pl->a.hy; return 0;
    case FDTD2D_JZ: rows = ny; cols = nx; *bufs = pl->a.jz; return 0;
  }
  return -1;
}

int fdtd2d_read_field(fdtd2d_plan *pl, int which, float *out, size_t count) {
  if (!pl || !out) return set_error(FDTD2D_EINVAL, "null argument");
  int rows, cols;
  float *const *bufs;
  if (field_shape(pl, which, rows, cols, &bufs)) return set_error(FDTD2D_EINVAL, "unknown field");
  if (count != static_cast<size_t>(rows) * cols) return set_error(FDTD2D_EINVAL, "field count mismatch");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
  HIP_TRY(hipMemcpy2D(out, cols * 4, bufs[pl->src], pl->L.pitch * 4, cols * 4, rows, hipMemcpyDeviceToHost),
          "hipMemcpy2D D2H field");
  return FDTD2D_OK;
}

int fdtd2d_write_field(fdtd2d_plan *pl, int which, const float *in, size_t count) {
  if (!pl || !in) return set_error(FDTD2D_EINVAL, "null argument");
  int rows, cols;
  float *const *bufs;
  if (field_shape(pl, which, rows, cols, &bufs)) return set_error(FDTD2D_EINVAL, "unknown field");
  if (count != static_cast<size_t>(rows) * cols) return set_error(FDTD2D_EINVAL, "field count mismatch");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
  // Both buffer sets hold the value: cells a step never updates must agree.
  for (int s = 0; s < 2; ++s)
    HIP_TRY(hipMemcpy2D(bufs[s], pl->L.pitch * 4, in, cols * 4, cols * 4, rows, hipMemcpyHostToDevice),
            "hipMemcpy2D H2D field");
  if (which == FDTD2D_JZ) {
    // tiles whose window now holds a nonzero Jz (any bits but +0) must update it
    for (int j = 0; j < rows; ++j) {
      int lo = INT32_MAX, hi = -1;
      for (int i = 0; i < cols; ++i) {
        uint32_t b;
        std::memcpy(&b, &in[static_cast<size_t>(j) * cols + i], 4);
        if (b != 0u) {
          lo = std::min(lo, i);
          hi = std::max(hi, i);
        }
      }
      pl->jz_lo[j] = lo;
      pl->jz_hi[j] = hi;
    }
    return set_tiling(pl);
  }
  return FDTD2D_OK;
}

static int psi_io(fdtd2d_plan *pl, int which, float *out, const float *in, size_t count) {
  const int P = pl->L.P, nx = pl->L.nx, ny = pl->L.ny, pitch = pl->L.pitch, xs = pl->L.xs;
  if (P == 0) return count == 0 ? FDTD2D_OK : set_error(FDTD2D_EINVAL, "no CPML in this plan");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(pl->stream), "hipStreamSynchronize");
  const bool ystrip = which == FDTD2D_PSI_HXY || which == FDTD2D_PSI_EZY;
  const int nr = (which == FDTD2D_PSI_HXY || which == FDTD2D_PSI_HYX) ? P - 1 : P;
  const size_t per_side = ystrip ? static_cast<size_t>(nr) * nx : static_cast<size_t>(ny) * nr;
  if (count != 2 * per_side) return set_error(FDTD2D_EINVAL, "psi count mismatch");
  float *dev[2];
  for (int s = 0; s < 2; ++s) {
    switch (which) {
      case FDTD2D_PSI_HXY: dev[s] = pl->a.psi_hxy[s]; break;
      case FDTD2D_PSI_EZY: dev[s] = pl->a.psi_ezy[s]; break;
      case FDTD2D_PSI_HYX: dev[s] = pl->a.psi_hyx[s]; break;
      case FDTD2D_PSI_EZX: dev[s] = pl->a.psi_ezx[s]; break;
      default: return set_error(FDTD2D_EINVAL, "unknown psi array");
    }
  }
  for (int side = 0; side < 2; ++side) {
    const size_t ho = side * per_side;
    if (ystrip) {  // device rows slot = side*P + r, r in [0, nr)
      const size_t doff = static_cast<size_t>(side * P) * pitch;
      if (out) {
        HIP_TRY(hipMemcpy2D(out + ho, nx * 4, dev[pl->src] + doff, pitch * 4, nx * 4, nr, hipMemcpyDeviceToHost),
                "hipMemcpy2D psi");
      } else {
        for (int s = 0; s < 2; ++s)
          HIP_TRY(hipMemcpy2D(dev[s] + doff, pitch * 4, in + ho, nx * 4, nx * 4, nr, hipMemcpyHostToDevice),
                  "hipMemcpy2D psi");
      }
    } else {  // device [ny][xs], column slot = side*P + r
      const size_t doff = static_cast<size_t>(side * P);
      if (out) {
        HIP_TRY(hipMemcpy2D(out + ho, nr * 4, dev[pl->src] + doff, xs * 4, nr * 4, ny, hipMemcpyDeviceToHost),
                "hipMemcpy2D psi");
      } else {
        for (int s = 0; s < 2; ++s)
          HIP_TRY(hipMemcpy2D(dev[s] + doff, xs * 4, in + ho, nr * 4, nr * 4, ny, hipMemcpyHostToDevice),
                  "hipMemcpy2D psi");
      }
    }
  }
  return FDTD2D_OK;
}

int fdtd2d_read_psi(fdtd2d_plan *pl, int which, float *out, size_t count) {
  if (!pl || !out) return set_error(FDTD2D_EINVAL, "null argument");
  return psi_io(pl, which, out, nullptr, count);
}

int fdtd2d_write_psi(fdtd2d_plan *pl, int which, const float *in, size_t count) {
  if (!pl || !in) return set_error(FDTD2D_EINVAL, "null argument");
  return psi_io(pl, which, nullptr, in, count);
}

int fdtd2d_plan_time_kernels(fdtd2d_plan *pl, int enable) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  int rc = fdtd2d_sync(pl);
  if (rc) return rc;
  pl->timing = enable != 0;
  for (int c = 0; c < 2; ++c) {
    pl->sum_ms[c] = 0;
    pl->count[c] = 0;
  }
  return FDTD2D_OK;
}

int fdtd2d_plan_kernel_time(fdtd2d_plan *pl, int cls, double *avg_ms, int *count) {
  if (!pl || !avg_ms || !count) return set_error(FDTD2D_EINVAL, "null argument");
  if (cls != 0 && cls != 1) return set_error(FDTD2D_EINVAL, "kernel class must be 0 or 1");
  int rc = fdtd2d_sync(pl);
  if (rc) return rc;
  *count = pl->count[cls];
  *avg_ms = pl->count[cls] ? pl->sum_ms[cls] / pl->count[cls] : 0.0;
  return FDTD2D_OK;
}

int fdtd2d_plan_kernel_ms(fdtd2d_plan *pl, double *avg_ms, int *count) {
  return fdtd2d_plan_kernel_time(pl, 0, avg_ms, count);
}

int fdtd2d_plan_destroy(fdtd2d_plan *pl) {
  if (!pl) return FDTD2D_OK;
  hipSetDevice(pl->device);
  if (pl->stream || pl->static_lists) hipStreamSynchronize(pl->stream);
  for (int c = 0; c < 2; ++c)
    for (auto &pr : pl->pending[c]) {
      hipEventDestroy(pr.first);
      hipEventDestroy(pr.second);
    }
  for (auto e : pl->spare) hipEventDestroy(e);
  if (pl->aux) hipStreamSynchronize(pl->aux);
  if (pl->ev_fork) hipEventDestroy(pl->ev_fork);
  if (pl->ev_join) hipEventDestroy(pl->ev_join);
  if (pl->holds_static) {
    std::lock_guard<std::mutex> lk(g_static_mu);
    g_static_busy[pl->device] = false;
  } else if (pl->wg_buf && !pl->static_lists) {
    hipFree(pl->wg_buf);
  }
  if (pl->base && pl->owns_base) hipFree(pl->base);
  if (pl->aux) hipStreamDestroy(pl->aux);
  if (pl->own_stream) hipStreamDestroy(pl->own_stream);
  delete pl;
  return FDTD2D_OK;
}

}  // extern "C"
