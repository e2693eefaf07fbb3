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
#include <new>
#include <string>
#include <vector>

#include "../../include/fdtd2d.h"
#include "errors.h"
#include "plan_internal.h"
#include "validate.h"

using fdtd2d::set_error;

namespace {

constexpr size_t kAlign = 256;
// Scratch rows below each field plane: interior tiles direct the stores of
// lanes/rows they do not own there, so their loop issues the same memory
// operations every row (nothing reads these rows).
constexpr int kTrashRows = 64;
// Steps per launch when none is set: on MI355X the interior kernel is issue
// bound near 8 levels and memory bound below (8192^2: 81 Gcell/s at 1,
// ~300 at 4, ~520 at 8 with 1 column per lane and 128-row bands).
constexpr int kDefaultT = 8;

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

struct Layout {
  int nx, ny, pitch, P, xs;
  size_t off_fields[2][4];
  size_t off_psi_y[2][2];  // [set][hxy, ezy]
  size_t off_psi_x[2][2];  // [set][hyx, ezx]
  size_t off_med, off_lut;
  size_t off_col[8];  // den_ex, den_hx, hyx_b, hyx_c, ezx_b, ezx_c, hyx_slot, ezx_slot
  size_t off_row[8];  // den_ey, den_hy, hxy_b, hxy_c, ezy_b, ezy_c, hxy_slot, ezy_slot
  size_t off_rowinfo, off_evoff, off_events, off_src, off_det, off_comp;
  size_t n_events;
  size_t total;
};

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
  L.off_rowinfo = take(static_cast<size_t>(L.ny) * 16);
  L.off_evoff = take(static_cast<size_t>(L.ny + 1) * 4);
  L.off_events = take(n_events * sizeof(fdtd2d::Event));
  L.off_src = take(static_cast<size_t>(p->nsteps + 1) * p->nsrc * 4);
  L.off_det = take(static_cast<size_t>(p->nsteps + 1) * p->ndet * 4);
  L.off_comp = take(static_cast<size_t>(3) * (p->nsteps + 1) * p->ncomp * 4);
  L.total = o;
  return L;
}

}  // namespace

struct fdtd2d_plan {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  char *base = nullptr;
  Layout L{};
  fdtd2d::StepArgs a{};
  int src = 0;
  int n = 0;
  int nsteps = 0;
  int T = 0;  // 0: default
  int C = 1;
  int band = 0;  // 0: automatic
  int D = 1;     // rows of prefetch per wave (interior kernel)
  int band_cpml = 0;               // rows per CPML-class work-group (0: automatic)
  int4 *wg_buf = nullptr;          // device copy of the two work-group lists
  hipStream_t aux = nullptr;       // CPML-class launches run here, beside the interior ones
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int ndet = 0, ncomp = 0, nsrc = 0;
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<hipEvent_t> spare;
  double sum_ms = 0;
  int count = 0;
  int launches_T[fdtd2d::kMaxT + 1] = {0};
};

namespace {

int hip_fail(hipError_t e, const char *what) {
  return set_error(FDTD2D_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                 \
  do {                                      \
    hipError_t e_ = (expr);                 \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

// Band height: enough bands for >= ~8 waves per SIMD on 256 CUs, at least 32
// rows, at most 256.
int pick_band(int nx, int ny, int VW, int T) {
  const int ntx = (nx + VW - 1) / VW;
  const long long target = 8LL * 1024;  // waves
  int H = 256;
  while (H > 32 && static_cast<long long>(ntx) * ((ny + H - 1) / H) < target) H /= 2;
  return std::max(H, T);
}

int set_tiling(fdtd2d_plan *pl) {
  const int T = pl->T > 0 ? pl->T : kDefaultT;
  const int halo = ((T + pl->C - 1) / pl->C) * pl->C;
  const int VW = 64 * pl->C - 2 * halo;
  fdtd2d::StepArgs &a = pl->a;
  a.halo = halo;
  a.vw = VW;
  a.ntx = (pl->L.nx + VW - 1) / VW;
  a.nwgx = (a.ntx + fdtd2d::kWavesPerWG - 1) / fdtd2d::kWavesPerWG;
  a.H = pl->band > 0 ? std::max(pl->band, T) : pick_band(pl->L.nx, pl->L.ny, VW, T);
  a.nty = (pl->L.ny + a.H - 1) / a.H;
  // work-group lists: interior / CPML, each in (column group, band) order;
  // CPML-class bands are cut into sub-bands of hc rows.
  const int hc = pl->band_cpml > 0 ? std::max(pl->band_cpml, T) : std::max(16, T);
  std::vector<int4> lists[2];
  for (int wgx = 0; wgx < a.nwgx; ++wgx)
    for (int ty = 0; ty < a.nty; ++ty) {
      const int jb0 = ty * a.H, jb1 = std::min(pl->L.ny, jb0 + a.H);
      if (!fdtd2d::wg_touches_cpml(a, T, wgx, jb0, jb1)) {
        lists[0].push_back(make_int4(wgx, jb0, jb1, 0));
        continue;
      }
      for (int r0 = jb0; r0 < jb1; r0 += hc) {
        const int r1 = std::min(jb1, r0 + hc);
        lists[fdtd2d::wg_touches_cpml(a, T, wgx, r0, r1) ? 1 : 0].push_back(make_int4(wgx, r0, r1, 0));
      }
    }
  const size_t n = lists[0].size() + lists[1].size();
  if (pl->wg_buf) {
    hipFree(pl->wg_buf);
    pl->wg_buf = nullptr;
  }
  HIP_TRY(hipMalloc(reinterpret_cast<void **>(&pl->wg_buf), std::max<size_t>(n, 1) * sizeof(int4)),
          "hipMalloc tile lists");
  HIP_TRY(hipMemcpy(pl->wg_buf, lists[0].data(), lists[0].size() * sizeof(int4), hipMemcpyHostToDevice),
          "hipMemcpy lists");
  HIP_TRY(hipMemcpy(pl->wg_buf + lists[0].size(), lists[1].data(), lists[1].size() * sizeof(int4),
                    hipMemcpyHostToDevice),
          "hipMemcpy lists");
  a.wg_list[0] = pl->wg_buf;
  a.wg_list[1] = pl->wg_buf + lists[0].size();
  a.n_wg[0] = static_cast<int>(lists[0].size());
  a.n_wg[1] = static_cast<int>(lists[1].size());
  return FDTD2D_OK;
}

int current_T(const fdtd2d_plan *pl) { return pl->T > 0 ? pl->T : kDefaultT; }

int upload(fdtd2d_plan *pl, size_t off, const void *src, size_t bytes) {
  if (bytes == 0) return FDTD2D_OK;
  HIP_TRY(hipMemcpy(pl->base + off, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
  return FDTD2D_OK;
}

int drain_timing(fdtd2d_plan *pl) {
  for (auto &pr : pl->pending) {
    float ms = 0;
    HIP_TRY(hipEventSynchronize(pr.second), "hipEventSynchronize");
    HIP_TRY(hipEventElapsedTime(&ms, pr.first, pr.second), "hipEventElapsedTime");
    pl->sum_ms += ms;
    pl->count += 1;
    pl->spare.push_back(pr.first);
    pl->spare.push_back(pr.second);
  }
  pl->pending.clear();
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

}  // namespace

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
  pl->L = make_layout(p, events.size());
  pl->nsteps = p->nsteps;
  pl->ndet = p->ndet;
  pl->ncomp = p->ncomp;
  pl->nsrc = p->nsrc;
  const Layout &L = pl->L;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(&pl->base), L.total);
  if (e != hipSuccess) {
    delete pl;
    return set_error(FDTD2D_ENOMEM, std::string("hipMalloc of ") + std::to_string(L.total) +
                                        " bytes: " + hipGetErrorString(e));
  }
  auto fail = [&](int code) {
    fdtd2d_plan_destroy(pl);
    return code;
  };
  if (hipStreamCreateWithFlags(&pl->own_stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&pl->aux, hipStreamNonBlocking) != hipSuccess)
    return fail(set_error(FDTD2D_EDEVICE, "hipStreamCreate"));
  if (hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail(set_error(FDTD2D_EDEVICE, "hipEventCreate"));
  pl->stream = pl->own_stream;
  if (hipMemset(pl->base, 0, L.total) != hipSuccess) return fail(set_error(FDTD2D_EDEVICE, "hipMemset"));

  const int nx = L.nx, ny = L.ny, pitch = L.pitch, P = L.P;
  // medium in pitch layout
  {
    std::vector<uint32_t> med(static_cast<size_t>(ny) * pitch, 0u);
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        size_t c = static_cast<size_t>(j) * nx + i;
        uint32_t v;
        if (p->medium_kind == FDTD2D_MEDIUM_TABLE) {
          v = static_cast<uint32_t>(p->material[c]);
        } else {
          std::memcpy(&v, &p->density[c], 4);
        }
        med[static_cast<size_t>(j) * pitch + i] = v;
      }
    if ((rc = upload(pl, L.off_med, med.data(), med.size() * 4))) return fail(rc);
  }
  if (p->medium_kind == FDTD2D_MEDIUM_TABLE) {
    std::vector<float> lut(static_cast<size_t>(p->nmaterials) * 4, 0.f);
    for (int m = 0; m < p->nmaterials; ++m) {
      lut[4 * m + 0] = p->ca[m];
      lut[4 * m + 1] = p->cb[m];
      lut[4 * m + 2] = p->ne[m];
    }
    if ((rc = upload(pl, L.off_lut, lut.data(), lut.size() * 4))) return fail(rc);
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
      if ((rc = upload(pl, L.off_col[f], col[f].data(), pitch * 4))) return fail(rc);
    if ((rc = upload(pl, L.off_col[6], hyx_s.data(), pitch * 4))) return fail(rc);
    if ((rc = upload(pl, L.off_col[7], ezx_s.data(), pitch * 4))) return fail(rc);
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
      if ((rc = upload(pl, L.off_row[f], row[f].data(), ny * 4))) return fail(rc);
    if ((rc = upload(pl, L.off_row[6], hxy_s.data(), ny * 4))) return fail(rc);
    if ((rc = upload(pl, L.off_row[7], ezy_s.data(), ny * 4))) return fail(rc);
  }
  if ((rc = upload(pl, L.off_evoff, ev_off.data(), ev_off.size() * 4))) return fail(rc);
  {
    std::vector<int4> ri(ny);
    for (int j = 0; j < ny; ++j) {
      const int jc = std::min(j, ny - 2);
      int dh, de;
      std::memcpy(&dh, &p->den_hy[jc], 4);
      std::memcpy(&de, &p->den_ey[jc], 4);
      ri[j] = make_int4(dh, de, ev_off[j], ev_off[j + 1]);
    }
    if ((rc = upload(pl, L.off_rowinfo, ri.data(), ri.size() * sizeof(int4)))) return fail(rc);
  }
  if ((rc = upload(pl, L.off_events, events.data(), events.size() * sizeof(fdtd2d::Event)))) return fail(rc);
  if ((rc = upload(pl, L.off_src, p->src_table, static_cast<size_t>(p->nsteps + 1) * p->nsrc * 4)))
    return fail(rc);

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
  a.jz_dt = p->jz_dt;
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
  a.rowinfo = reinterpret_cast<const int4 *>(pl->base + L.off_rowinfo);
  a.events = reinterpret_cast<const fdtd2d::Event *>(pl->base + L.off_events);
  a.src_table = reinterpret_cast<const float *>(pl->base + L.off_src);
  a.nsrc = p->nsrc;
  a.det = reinterpret_cast<float *>(pl->base + L.off_det);
  a.ndet = p->ndet;
  a.comp = reinterpret_cast<float *>(pl->base + L.off_comp);
  a.ncomp = p->ncomp;
  a.comp_plane = static_cast<long long>(p->nsteps + 1) * p->ncomp;
  if ((rc = set_tiling(pl))) return fail(rc);
  *out = pl;
  return FDTD2D_OK;
}

int fdtd2d_plan_set_stream(fdtd2d_plan *pl, void *stream) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  pl->stream = stream ? static_cast<hipStream_t>(stream) : pl->own_stream;
  return FDTD2D_OK;
}

int fdtd2d_plan_set_tblock(fdtd2d_plan *pl, int T) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  if (T < 0 || T > fdtd2d::kMaxT) return set_error(FDTD2D_EINVAL, "steps per launch must be 0..8");
  pl->T = T;
  return set_tiling(pl);
}

int fdtd2d_plan_set_option(fdtd2d_plan *pl, const char *name, double value) {
  if (!pl || !name) return set_error(FDTD2D_EINVAL, "null argument");
  std::string n(name);
  const int v = static_cast<int>(value);
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
  } else if (n == "prefetch") {
    const int maxd = pl->C == 1 ? 4 : pl->C == 2 ? 2 : 1;
    if (v != 1 && v != 2 && v != 4) return set_error(FDTD2D_EINVAL, "prefetch must be 1, 2 or 4");
    if (v > maxd) return set_error(FDTD2D_EINVAL, "prefetch depth too large for cols_per_lane");
    pl->D = v;
  } else {
    return set_error(FDTD2D_EINVAL, "unknown option " + n);
  }
  return set_tiling(pl);
}

int fdtd2d_plan_get_info(const fdtd2d_plan *pl, const char *name, double *value) {
  if (!pl || !name || !value) return set_error(FDTD2D_EINVAL, "null argument");
  std::string n(name);
  if (n == "pitch") *value = pl->L.pitch;
  else if (n == "tblock") *value = current_T(pl);
  else if (n == "band") *value = pl->a.H;
  else if (n == "ntx") *value = pl->a.ntx;
  else if (n == "nty") *value = pl->a.nty;
  else if (n == "workgroups") *value = pl->a.n_wg[0] + pl->a.n_wg[1];
  else if (n == "workgroups_cpml") *value = pl->a.n_wg[1];
  else if (n == "tile_cols") *value = pl->a.vw;
  else if (n == "bytes") *value = static_cast<double>(pl->L.total);
  else if (n == "steps") *value = pl->n;
  else if (n == "cols_per_lane") *value = pl->C;
  else if (n == "prefetch") *value = pl->D;
  else return set_error(FDTD2D_EINVAL, "unknown info name " + n);
  return FDTD2D_OK;
}

int fdtd2d_run(fdtd2d_plan *pl, int nsteps) {
  if (!pl) return set_error(FDTD2D_EINVAL, "null plan");
  if (nsteps < 0 || pl->n + nsteps > pl->nsteps)
    return set_error(FDTD2D_EINVAL, "run beyond the problem's nsteps (" + std::to_string(pl->nsteps) + ")");
  HIP_TRY(hipSetDevice(pl->device), "hipSetDevice");
  const int Tmax = current_T(pl);
  while (nsteps > 0) {
    const int t = std::min(Tmax, nsteps);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (pl->timing) {
      e0 = get_event(pl);
      e1 = get_event(pl);
      if (!e0 || !e1) return set_error(FDTD2D_EDEVICE, "hipEventCreate");
      HIP_TRY(hipEventRecord(e0, pl->stream), "hipEventRecord");
    }
    // CPML work-groups on the aux stream, interior ones on the main stream;
    // both read buffer set src and write the other, so they may overlap.
    if (pl->a.n_wg[1] > 0) {
      HIP_TRY(hipEventRecord(pl->ev_fork, pl->stream), "hipEventRecord fork");
      HIP_TRY(hipStreamWaitEvent(pl->aux, pl->ev_fork, 0), "hipStreamWaitEvent fork");
      hipError_t e = fdtd2d::launch_step(pl->a, 1, t, pl->C, 1, pl->src, pl->n, pl->aux);
      if (e != hipSuccess) return hip_fail(e, "CPML step kernel launch");
      HIP_TRY(hipEventRecord(pl->ev_join, pl->aux), "hipEventRecord join");
    }
    hipError_t e = fdtd2d::launch_step(pl->a, 0, t, pl->C, pl->D, pl->src, pl->n, pl->stream);
    if (e != hipSuccess) return hip_fail(e, "step kernel launch");
    if (pl->a.n_wg[1] > 0) HIP_TRY(hipStreamWaitEvent(pl->stream, pl->ev_join, 0), "hipStreamWaitEvent join");
    if (pl->timing) {
      HIP_TRY(hipEventRecord(e1, pl->stream), "hipEventRecord");
      pl->pending.push_back({e0, e1});
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
  pl->sum_ms = 0;
  pl->count = 0;
  return FDTD2D_OK;
}

int fdtd2d_plan_kernel_ms(fdtd2d_plan *pl, double *avg_ms, int *count) {
  if (!pl || !avg_ms || !count) return set_error(FDTD2D_EINVAL, "null argument");
  int rc = fdtd2d_sync(pl);
  if (rc) return rc;
  *count = pl->count;
  *avg_ms = pl->count ? pl->sum_ms / pl->count : 0.0;
  return FDTD2D_OK;
}

int fdtd2d_plan_destroy(fdtd2d_plan *pl) {
  if (!pl) return FDTD2D_OK;
  hipSetDevice(pl->device);
  if (pl->stream) hipStreamSynchronize(pl->stream);
  for (auto &pr : pl->pending) {
    hipEventDestroy(pr.first);
    hipEventDestroy(pr.second);
  }
  for (auto e : pl->spare) hipEventDestroy(e);
  if (pl->aux) hipStreamSynchronize(pl->aux);
  if (pl->ev_fork) hipEventDestroy(pl->ev_fork);
  if (pl->ev_join) hipEventDestroy(pl->ev_join);
  if (pl->wg_buf) hipFree(pl->wg_buf);
  if (pl->base) hipFree(pl->base);
  if (pl->aux) hipStreamDestroy(pl->aux);
  if (pl->own_stream) hipStreamDestroy(pl->own_stream);
  delete pl;
  return FDTD2D_OK;
}

}  // extern "C"
