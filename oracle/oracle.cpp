// oracle.cpp - single-threaded CPU restatement of the reference's 2D TMz step.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP stepper. Built with
// -ffp-contract=off so that every float operation rounds exactly as written.
//
// Each loop below follows one reference kernel, with the kernel's own index
// ranges and operation order (FDTD_2D_Full/kernel.cu, abbreviated F/kernel.cu):
//   hx_main   F/kernel.cu:167-196   Dev_HxCalc
//   hx_pml    F/kernel.cu:198-240   Dev_HxPML_Jdir_Calc
//   hy_main   F/kernel.cu:288-318   Dev_HyCalc
//   hy_pml    F/kernel.cu:321-362   Dev_HyPML_Idir_Calc
//   jz        F/kernel.cu:127-163   Dev_JzCalc (E^n, FP64 promotion via double dt)
//   ez_main   F/kernel.cu:413-451   Dev_EzCalc (PEC id -> 0)
//   ez_pml_x  F/kernel.cu:453-494   Dev_EzPML_Idir_Calc
//   ez_pml_y  F/kernel.cu:496-537   Dev_EzPML_Jdir_Calc
//   source    F/kernel.cu:621-659   Dev_AESASourceCalc (value precomputed per step)
//   detectors F/kernel.cu:762-805   Dev_AESADetectorSim
//   comps     F/kernel.cu:719-759   Dev_AESAComponentDetector
//   rc lines  F/kernel.cu:700-713, 968-1044   Dev_FindMax, Dev_RCcalc
// and the step order of the host loop F/kernel.cu:2065-2157.
//
// What pins each function independently of this file (tests/test_oracle_kat.py
// unless noted), chosen so that a dropped term, a wrong sign or a shifted
// index fails at least one:
//   hx_main, hy_main, ez_main  discrete PEC-cavity eigenmode against the Yee
//                              dispersion closed form (FP64 to 1e-10), a
//                              perturbed operator failing it, discrete energy
//                              conservation, reciprocity in an inhomogeneous box
//   jz                         uniform-plasma modal recurrence (2x2 closed form,
//                              collisionless and collisional), plasma cutoff
//   hx_pml .. ez_pml_y         CPML reflection below -40 dB against a large
//                              domain, with a planted psi mistake failing it;
//                              coefficient closed forms (tests/test_setup.py)
//   ez_main PEC id             PEC material forces Ez = 0
//   source, detectors, comps   marked-grid brute force of every index rule
//   rc lines                   step-by-step brute force of Dev_RCcalc (a
//                              planted column mistake fails it); the full
//                              33568-step run of the reference's own
//                              Settings.ini reproduces its RC_Line.txt and
//                              TAU_Line.txt line for line
//   whole step                 FP32 within tolerance of FP64; threads bitwise
// No function here is left without such a check. Left unpinned, by nature:
// the reference's sm_20 FMA contraction, CUDA 7.5 device cos/pow/exp in its
// disabled chirp path, and its undefined outputs (out-of-bounds trace row,
// uninitialised buffers); these are defined here as documented above.
//
// Departures, all deliberate:
//  * The reference launches truncated grids (F/kernel.cu:2028-2030) that can
//    skip psi cells for some sizes; the loops here cover the kernels' full
//    stated ranges (the launch-size artefact is not reproduced).
//  * The reference writes trace row n for n = 1..nMax into an nMax-row buffer
//    (row nMax out of bounds); traces here have nsteps+1 rows, row 0 = 0.
//  * The reference H tables DA/DB are per material but equal 1 and dt/mu0 for
//    every material it can build (F/FDTD_cuda.cu:521-533, 1196-1199): the
//    problem carries the single value db, and DA = 1 is applied literally.
//  * The Jz collision extension: jz_decay = 1, jz_dt = dt reproduces
//    F/kernel.cu:155 bit for bit.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

struct SimBase {
  virtual ~SimBase() {}
  virtual int run(int nsteps) = 0;
  virtual int read_field(int which, double *out, size_t count) const = 0;
  virtual int write_field(int which, const double *in, size_t count) = 0;
  virtual int read_psi(int which, double *out, size_t count) const = 0;
  virtual int write_psi(int which, const double *in, size_t count) = 0;
  virtual int read_det(double *det, size_t dc, double *comp, size_t cc) const = 0;
  virtual int read_rc(int32_t *spd, size_t count) const = 0;
  virtual int read_line(double *out, size_t count) const = 0;
  int nthreads = 1;
  int n = 0;
};

template <class V>
void vcopy(std::vector<V> &dst, const float *src, size_t count) {
  dst.assign(count, V(0));
  if (src)
    for (size_t k = 0; k < count; ++k) dst[k] = V(src[k]);
}

// Row-parallel sweep helper: rows [j0, j1) split across nthreads.
template <class F>
void rows(int nthreads, int j0, int j1, F &&f) {
  if (j1 <= j0) return;
  if (nthreads <= 1 || j1 - j0 < 64) {
    for (int j = j0; j < j1; ++j) f(j);
    return;
  }
  std::vector<std::thread> th;
  int nt = nthreads;
  int chunk = (j1 - j0 + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    int a = j0 + t * chunk, b = std::min(j1, a + chunk);
    if (a >= b) break;
    th.emplace_back([=, &f]() {
      for (int j = a; j < b; ++j) f(j);
    });
  }
  for (auto &x : th) x.join();
}

// R = float reproduces the reference's FP32 path; R = double is the FP64 build.
template <class R>
struct Sim : SimBase {
  fdtd2d_problem p;
  int nx, ny, P;
  std::vector<R> Ez, Hx, Hy, Jz;
  std::vector<R> psi_hxy[2], psi_hyx[2], psi_ezx[2], psi_ezy[2];
  std::vector<R> be_x[2], ce_x[2], bh_x[2], ch_x[2];
  std::vector<R> be_y[2], ce_y[2], bh_y[2], ch_y[2];
  std::vector<R> den_ex, den_hx, den_ey, den_hy;
  std::vector<R> ca, cb, ne, density;
  std::vector<int32_t> material;
  std::vector<R> src_table;
  std::vector<int32_t> src_i, src_j, det_i, det_j, comp_i, comp_j;
  std::vector<R> det, comp;
  std::vector<R> line;  // processing line captures [rows][line_size]
  int line_rows = 0;
  std::vector<int32_t> rc_i, rc_imax, spd;
  std::vector<R> rc_edge;
  R rc_level;
  R da, db, ca0, cb0, cj, jz_decay, dxr, dyr;
  double jz_dt;

  explicit Sim(const fdtd2d_problem &q) : p(q) {
    nx = q.nx;
    ny = q.ny;
    P = q.npml;
    Ez.assign((size_t)nx * ny, R(0));
    Jz.assign((size_t)nx * ny, R(0));
    Hx.assign((size_t)nx * (ny - 1), R(0));
    Hy.assign((size_t)(nx - 1) * ny, R(0));
    for (int s = 0; s < 2; ++s) {
      if (P > 0) {
        psi_hxy[s].assign((size_t)(P - 1) * nx, R(0));
        psi_hyx[s].assign((size_t)ny * (P - 1), R(0));
        psi_ezx[s].assign((size_t)ny * P, R(0));
        psi_ezy[s].assign((size_t)P * nx, R(0));
        vcopy(be_x[s], q.be_x[s], P);
        vcopy(ce_x[s], q.ce_x[s], P);
        vcopy(bh_x[s], q.bh_x[s], P - 1);
        vcopy(ch_x[s], q.ch_x[s], P - 1);
        vcopy(be_y[s], q.be_y[s], P);
        vcopy(ce_y[s], q.ce_y[s], P);
        vcopy(bh_y[s], q.bh_y[s], P - 1);
        vcopy(ch_y[s], q.ch_y[s], P - 1);
      }
    }
    vcopy(den_ex, q.den_ex, nx - 1);
    vcopy(den_hx, q.den_hx, nx - 1);
    vcopy(den_ey, q.den_ey, ny - 1);
    vcopy(den_hy, q.den_hy, ny - 1);
    if (q.medium_kind == FDTD2D_MEDIUM_TABLE) {
      material.assign(q.material, q.material + (size_t)nx * ny);
      vcopy(ca, q.ca, q.nmaterials);
      vcopy(cb, q.cb, q.nmaterials);
      vcopy(ne, q.ne, q.nmaterials);
    } else {
      vcopy(density, q.density, (size_t)nx * ny);
    }
    vcopy(src_table, q.src_table, (size_t)(q.nsteps + 1) * q.nsrc);
    src_i.assign(q.src_i, q.src_i + q.nsrc);
    src_j.assign(q.src_j, q.src_j + q.nsrc);
    det_i.assign(q.det_i, q.det_i + q.ndet);
    det_j.assign(q.det_j, q.det_j + q.ndet);
    comp_i.assign(q.comp_i, q.comp_i + q.ncomp);
    comp_j.assign(q.comp_j, q.comp_j + q.ncomp);
    if (q.nrc > 0) {
      rc_i.assign(q.rc_i, q.rc_i + q.nrc);
      rc_imax.assign(q.rc_imax, q.rc_imax + q.nrc);
    }
    rc_edge.assign(q.nrc, R(q.rc_edge0));
    rc_level = R(q.rc_level);
    spd.assign((size_t)(q.nsteps + 1) * q.nrc, 0);
    line_rows = q.line_size > 0 ? std::max(0, std::min(q.line_end, q.nsteps) - q.line_begin) : 0;
    line.assign((size_t)line_rows * (line_rows ? q.line_size : 0), R(0));
    det.assign((size_t)(q.nsteps + 1) * q.ndet, R(0));
    comp.assign((size_t)3 * (q.nsteps + 1) * q.ncomp, R(0));
    da = R(1);
    db = R(q.db);
    ca0 = R(q.ca0);
    cb0 = R(q.cb0);
    cj = R(q.cj);
    jz_decay = R(q.jz_decay);
    jz_dt = q.jz_dt;
    // The reference passes dx, dy to the psi kernels as float arguments.
    dxr = sizeof(R) == sizeof(float) ? R(q.dx_f) : R(q.dx);
    dyr = sizeof(R) == sizeof(float) ? R(q.dy_f) : R(q.dy);
  }

  // Jz += CJ*Ez*Ne*dt: the float product is promoted to double by the double
  // dt argument and the sum rounded back to float (F/kernel.cu:127, 155).
  static inline R jz_update(R jz, R decay, R cjv, R ez, R nev, double kdt) {
    if (sizeof(R) == sizeof(float)) {
      R a = decay * jz;
      R b = (cjv * ez) * nev;
      return R(double(a) + double(b) * kdt);
    } else {
      return decay * jz + ((cjv * ez) * nev) * R(kdt);
    }
  }

  void step() {
    const int n_step = n + 1;
    const size_t NX = (size_t)nx;
    const size_t NH = (size_t)(nx - 1);  // Hy pitch
    // ---- Hx main (F/kernel.cu:167-196): i in [0,nx-2], j in [0,ny-2]
    rows(nthreads, 0, ny - 1, [&](int j) {
      for (int i = 0; i <= nx - 2; ++i) {
        size_t c = j * NX + i;
        Hx[c] = da * Hx[c] + db * ((Ez[c] - Ez[c + NX]) * den_hy[j]);
      }
    });
    // ---- Hx y-CPML (F/kernel.cu:198-240): j in [0,P-2], mirror jj = ny-2-j
    if (P > 1) {
      for (int j = 0; j <= P - 2; ++j)
        for (int i = 0; i <= nx - 2; ++i) {
          size_t c = j * NX + i;
          R &ps = psi_hxy[0][j * NX + i];
          ps = bh_y[0][j] * ps + ch_y[0][j] * (Ez[c] - Ez[c + NX]) / dyr;
          Hx[c] = Hx[c] + db * ps;
          int jj = (ny - 2) - j;
          size_t c2 = jj * NX + i;
          R &ps2 = psi_hxy[1][j * NX + i];
          ps2 = bh_y[1][j] * ps2 + ch_y[1][j] * (Ez[c2] - Ez[c2 + NX]) / dyr;
          Hx[c2] = Hx[c2] + db * ps2;
        }
    }
    // ---- Hy main (F/kernel.cu:288-318): i in [0,nx-2], j in [0,ny-2], pitch nx-1
    rows(nthreads, 0, ny - 1, [&](int j) {
      for (int i = 0; i <= nx - 2; ++i) {
        size_t h = j * NH + i, e = j * NX + i;
        Hy[h] = da * Hy[h] + db * ((Ez[e + 1] - Ez[e]) * den_hx[i]);
      }
    });
    // ---- Hy x-CPML (F/kernel.cu:321-362): i in [0,P-2], mirror ii = nx-2-i
    if (P > 1) {
      rows(nthreads, 0, ny - 1, [&](int j) {
        for (int i = 0; i <= P - 2; ++i) {
          size_t q = j * (size_t)(P - 1) + i;
          size_t e = j * NX + i, h = j * NH + i;
          R &ps = psi_hyx[0][q];
          ps = bh_x[0][i] * ps + ch_x[0][i] * (Ez[e + 1] - Ez[e]) / dxr;
          Hy[h] = Hy[h] + db * ps;
          int ii = (nx - 2) - i;
          size_t e2 = j * NX + ii, h2 = j * NH + ii;
          R &ps2 = psi_hyx[1][q];
          ps2 = bh_x[1][i] * ps2 + ch_x[1][i] * (Ez[e2 + 1] - Ez[e2]) / dxr;
          Hy[h2] = Hy[h2] + db * ps2;
        }
      });
    }
    // ---- Jz (F/kernel.cu:127-163): i,j in [1,N-2], where Ne > 0, from E^n
    const bool table = p.medium_kind == FDTD2D_MEDIUM_TABLE;
    rows(nthreads, 1, ny - 1, [&](int j) {
      for (int i = 1; i <= nx - 2; ++i) {
        size_t c = j * NX + i;
        R nev = table ? ne[material[c]] : density[c];
        if (nev > R(0)) Jz[c] = jz_update(Jz[c], jz_decay, cj, Ez[c], nev, jz_dt);
      }
    });
    // ---- Ez main (F/kernel.cu:413-451): i,j in [1,N-2]
    rows(nthreads, 1, ny - 1, [&](int j) {
      for (int i = 1; i <= nx - 2; ++i) {
        size_t c = j * NX + i, h = j * NH + i;
        R cav, cbv;
        if (table) {
          int id = material[c];
          if (id == p.pec_material) {
            Ez[c] = R(0);
            continue;
          }
          cav = ca[id];
          cbv = cb[id];
        } else {
          cav = ca0;
          cbv = cb0;
        }
        Ez[c] = cav * Ez[c] +
                cbv * ((Hy[h] - Hy[h - 1]) * den_ex[i] + (Hx[c - NX] - Hx[c]) * den_ey[j] - Jz[c]);
      }
    });
    auto cb_at = [&](size_t c) -> R { return table ? cb[material[c]] : cb0; };
    // ---- Ez x-CPML (F/kernel.cu:453-494): i in [1,P-1], mirror ii = nx-1-i, j in [1,ny-2]
    if (P > 1) {
      rows(nthreads, 1, ny - 1, [&](int j) {
        for (int i = 1; i <= P - 1; ++i) {
          size_t q = j * (size_t)P + i;
          size_t c = j * NX + i, h = j * NH + i;
          R &ps = psi_ezx[0][q];
          ps = be_x[0][i] * ps + ce_x[0][i] * (Hy[h] - Hy[h - 1]) / dxr;
          Ez[c] = Ez[c] + cb_at(c) * ps;
          int ii = (nx - 1) - i;
          size_t c2 = j * NX + ii, h2 = j * NH + ii;
          R &ps2 = psi_ezx[1][q];
          ps2 = be_x[1][i] * ps2 + ce_x[1][i] * (Hy[h2] - Hy[h2 - 1]) / dxr;
          Ez[c2] = Ez[c2] + cb_at(c2) * ps2;
        }
      });
      // ---- Ez y-CPML (F/kernel.cu:496-537): j in [1,P-1], mirror jj = ny-1-j, i in [1,nx-2]
      for (int j = 1; j <= P - 1; ++j)
        for (int i = 1; i <= nx - 2; ++i) {
          size_t c = j * NX + i;
          R &ps = psi_ezy[0][j * NX + i];
          ps = be_y[0][j] * ps + ce_y[0][j] * (Hx[c - NX] - Hx[c]) / dyr;
          Ez[c] = Ez[c] + cb_at(c) * ps;
          int jj = (ny - 1) - j;
          size_t c2 = jj * NX + i;
          R &ps2 = psi_ezy[1][j * NX + i];
          ps2 = be_y[1][j] * ps2 + ce_y[1][j] * (Hx[c2 - NX] - Hx[c2]) / dyr;
          Ez[c2] = Ez[c2] + cb_at(c2) * ps2;
        }
    }
    // ---- soft sources (F/kernel.cu:621-659), element order
    for (int e = 0; e < p.nsrc; ++e) {
      size_t c = src_j[e] * NX + src_i[e];
      Ez[c] = Ez[c] - src_table[(size_t)n_step * p.nsrc + e];
    }
    // ---- processing line (Dev_SaveLine F/kernel.cu:666-688, copied out at
    // :2138-2145 for begin_after < n <= finish_at): Ez of row line_j, columns
    // [line_i0, line_i0 + line_size), the line the kernel's arguments name
    if (line_rows > 0 && n_step > p.line_begin && n_step - p.line_begin <= line_rows) {
      const size_t t = (size_t)(n_step - p.line_begin - 1);
      for (int k = 0; k < p.line_size; ++k) line[t * p.line_size + k] = Ez[p.line_j * NX + p.line_i0 + k];
    }
    // ---- detectors (F/kernel.cu:762-805) and components (F/kernel.cu:719-759)
    for (int d = 0; d < p.ndet; ++d)
      det[(size_t)n_step * p.ndet + d] = Ez[det_j[d] * NX + det_i[d]];
    const size_t plane = (size_t)(p.nsteps + 1) * p.ncomp;
    for (int c = 0; c < p.ncomp; ++c) {
      size_t row = (size_t)n_step * p.ncomp + c;
      comp[row] = Ez[comp_j[c] * NX + comp_i[c]];
      comp[plane + row] = Hx[comp_j[c] * NX + comp_i[c]];
      comp[2 * plane + row] = Hy[comp_j[c] * NH + comp_i[c]];
    }
    // ---- RC cutoff lines, after the detectors (F/kernel.cu:2170): Dev_RCcalc
    // (:968-1044) with Dev_FindMax (:700-713). The maximum is signed and seeded
    // from row 0 of the maximum column; the scan starts below row ny-2.
    for (int e = 0; e < p.nrc; ++e) {
      const size_t im = (size_t)rc_imax[e], is = (size_t)rc_i[e];
      R temp = Ez[im];
      for (int j = p.rc_jcut; j < ny - 2; ++j)
        if (Ez[j * NX + im] > temp) temp = Ez[j * NX + im];
      const R edge_tmp = temp * rc_level;
      if (edge_tmp > rc_edge[e]) rc_edge[e] = edge_tmp;
      int j = ny - 2;
      do {
        j--;
      } while (j > 0 && std::abs(Ez[j * NX + is]) < rc_edge[e]);
      spd[(size_t)n_step * p.nrc + e] = j;
    }
    n = n_step;
  }

  int run(int nsteps) override {
    if (nsteps < 0 || n + nsteps > p.nsteps)
      return fail(FDTD2D_EINVAL, "oracle_run: steps beyond the problem's nsteps");
    for (int k = 0; k < nsteps; ++k) step();
    return FDTD2D_OK;
  }

  std::vector<R> *field(int which) {
    switch (which) {
      case FDTD2D_EZ: return &Ez;
      case FDTD2D_HX: return &Hx;
      case FDTD2D_HY: return &Hy;
      case FDTD2D_JZ: return &Jz;
    }
    return nullptr;
  }
  int read_field(int which, double *out, size_t count) const override {
    auto *f = const_cast<Sim *>(this)->field(which);
    if (!f || count != f->size()) return fail(FDTD2D_EINVAL, "oracle_read_field: bad field/count");
    for (size_t k = 0; k < count; ++k) out[k] = double((*f)[k]);
    return FDTD2D_OK;
  }
  int write_field(int which, const double *in, size_t count) override {
    auto *f = field(which);
    if (!f || count != f->size()) return fail(FDTD2D_EINVAL, "oracle_write_field: bad field/count");
    for (size_t k = 0; k < count; ++k) (*f)[k] = R(in[k]);
    return FDTD2D_OK;
  }
  std::vector<R> *psi(int which) {
    switch (which) {
      case FDTD2D_PSI_HXY: return psi_hxy;
      case FDTD2D_PSI_HYX: return psi_hyx;
      case FDTD2D_PSI_EZX: return psi_ezx;
      case FDTD2D_PSI_EZY: return psi_ezy;
    }
    return nullptr;
  }
  int read_psi(int which, double *out, size_t count) const override {
    auto *f = const_cast<Sim *>(this)->psi(which);
    if (!f || count != f[0].size() + f[1].size())
      return fail(FDTD2D_EINVAL, "oracle_read_psi: bad strip/count");
    size_t k = 0;
    for (int s = 0; s < 2; ++s)
      for (auto v : f[s]) out[k++] = double(v);
    return FDTD2D_OK;
  }
  int write_psi(int which, const double *in, size_t count) override {
    auto *f = psi(which);
    if (!f || count != f[0].size() + f[1].size())
      return fail(FDTD2D_EINVAL, "oracle_write_psi: bad strip/count");
    size_t k = 0;
    for (int s = 0; s < 2; ++s)
      for (auto &v : f[s]) v = R(in[k++]);
    return FDTD2D_OK;
  }
  int read_line(double *out, size_t count) const override {
    if (count != line.size()) return fail(FDTD2D_EINVAL, "oracle_read_line: count must be rows*line_size");
    for (size_t k = 0; k < count; ++k) out[k] = double(line[k]);
    return FDTD2D_OK;
  }
  int read_rc(int32_t *out, size_t count) const override {
    if (count != spd.size()) return fail(FDTD2D_EINVAL, "oracle_read_rc: count must be (nsteps+1)*nrc");
    for (size_t k = 0; k < count; ++k) out[k] = spd[k];
    return FDTD2D_OK;
  }
  int read_det(double *d, size_t dc, double *c, size_t cc) const override {
    if (d) {
      if (dc != det.size()) return fail(FDTD2D_EINVAL, "oracle_read_detectors: det count");
      for (size_t k = 0; k < dc; ++k) d[k] = double(det[k]);
    }
    if (c) {
      if (cc != comp.size()) return fail(FDTD2D_EINVAL, "oracle_read_detectors: comp count");
      for (size_t k = 0; k < cc; ++k) c[k] = double(comp[k]);
    }
    return FDTD2D_OK;
  }
};

int validate(const fdtd2d_problem *p) {
  if (!p) return fail(FDTD2D_EINVAL, "null problem");
  if (p->npml < 0 || p->npml == 1) return fail(FDTD2D_EINVAL, "npml must be 0 or >= 2");
  if (p->nx < 2 * p->npml + 4 || p->ny < 2 * p->npml + 4)
    return fail(FDTD2D_EINVAL, "grid smaller than 2*npml+4");
  if (p->nsteps < 0) return fail(FDTD2D_EINVAL, "nsteps < 0");
  if (!p->den_ex || !p->den_hx || !p->den_ey || !p->den_hy)
    return fail(FDTD2D_EINVAL, "missing den tables");
  if (p->npml > 0)
    for (int s = 0; s < 2; ++s)
      if (!p->be_x[s] || !p->ce_x[s] || !p->bh_x[s] || !p->ch_x[s] || !p->be_y[s] ||
          !p->ce_y[s] || !p->bh_y[s] || !p->ch_y[s])
        return fail(FDTD2D_EINVAL, "missing CPML tables");
  if (p->medium_kind == FDTD2D_MEDIUM_TABLE) {
    if (!p->material || !p->ca || !p->cb || !p->ne || p->nmaterials <= 0)
      return fail(FDTD2D_EINVAL, "TABLE medium needs material, ca, cb, ne");
    for (size_t k = 0; k < (size_t)p->nx * p->ny; ++k)
      if (p->material[k] < 0 || p->material[k] >= p->nmaterials)
        return fail(FDTD2D_EINVAL, "material id out of range");
  } else if (p->medium_kind == FDTD2D_MEDIUM_DENSITY) {
    if (!p->density) return fail(FDTD2D_EINVAL, "DENSITY medium needs density");
  } else {
    return fail(FDTD2D_EINVAL, "unknown medium_kind");
  }
  if (p->nsrc > 0 && (!p->src_i || !p->src_j || !p->src_table))
    return fail(FDTD2D_EINVAL, "sources need src_i, src_j, src_table");
  for (int e = 0; e < p->nsrc; ++e)
    if (p->src_i[e] < 0 || p->src_i[e] >= p->nx || p->src_j[e] < 0 || p->src_j[e] >= p->ny)
      return fail(FDTD2D_EINVAL, "source outside the grid");
  for (int d = 0; d < p->ndet; ++d)
    if (p->det_i[d] < 0 || p->det_i[d] >= p->nx || p->det_j[d] < 0 || p->det_j[d] >= p->ny)
      return fail(FDTD2D_EINVAL, "detector outside the grid");
  for (int c = 0; c < p->ncomp; ++c)
    if (p->comp_i[c] < 0 || p->comp_i[c] > p->nx - 2 || p->comp_j[c] < 0 ||
        p->comp_j[c] > p->ny - 2)
      return fail(FDTD2D_EINVAL, "component detector outside [0,nx-2]x[0,ny-2]");
  if (p->nrc < 0 || (p->nrc > 0 && (!p->rc_i || !p->rc_imax))) return fail(FDTD2D_EINVAL, "bad RC lines");
  for (int e = 0; e < p->nrc; ++e)
    if (p->rc_i[e] < 0 || p->rc_i[e] >= p->nx || p->rc_imax[e] < 0 || p->rc_imax[e] >= p->nx)
      return fail(FDTD2D_EINVAL, "RC line outside the grid");
  if (p->nrc > 0 && p->rc_jcut < 0) return fail(FDTD2D_EINVAL, "rc_jcut < 0");
  if (p->line_size < 0) return fail(FDTD2D_EINVAL, "negative processing line size");
  if (p->line_size > 0 && (p->line_i0 < 0 || (long long)p->line_i0 + p->line_size > p->nx || p->line_j < 0 ||
                           p->line_j >= p->ny || p->line_begin < 0))
    return fail(FDTD2D_EINVAL, "processing line outside the grid");
  return FDTD2D_OK;
}

}  // namespace

struct oracle_sim {
  SimBase *impl;
};

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

int oracle_create(const fdtd2d_problem *p, int precision, oracle_sim **out) {
  if (!out) return fail(FDTD2D_EINVAL, "null out");
  int rc = validate(p);
  if (rc) return rc;
  SimBase *impl = nullptr;
  if (precision == 32)
    impl = new Sim<float>(*p);
  else if (precision == 64)
    impl = new Sim<double>(*p);
  else
    return fail(FDTD2D_EINVAL, "precision must be 32 or 64");
  *out = new oracle_sim{impl};
  return FDTD2D_OK;
}

int oracle_set_threads(oracle_sim *s, int nthreads) {
  if (!s || nthreads < 1) return fail(FDTD2D_EINVAL, "bad thread count");
  s->impl->nthreads = nthreads;
  return FDTD2D_OK;
}

int oracle_run(oracle_sim *s, int nsteps) { return s ? s->impl->run(nsteps) : FDTD2D_EINVAL; }
int oracle_steps_done(const oracle_sim *s) { return s ? s->impl->n : -1; }
int oracle_read_field(const oracle_sim *s, int which, double *out, size_t count) {
  return s ? s->impl->read_field(which, out, count) : FDTD2D_EINVAL;
}
int oracle_write_field(oracle_sim *s, int which, const double *in, size_t count) {
  return s ? s->impl->write_field(which, in, count) : FDTD2D_EINVAL;
}
int oracle_read_psi(const oracle_sim *s, int which, double *out, size_t count) {
  return s ? s->impl->read_psi(which, out, count) : FDTD2D_EINVAL;
}
int oracle_write_psi(oracle_sim *s, int which, const double *in, size_t count) {
  return s ? s->impl->write_psi(which, in, count) : FDTD2D_EINVAL;
}
int oracle_read_detectors(const oracle_sim *s, double *det, size_t dc, double *comp, size_t cc) {
  return s ? s->impl->read_det(det, dc, comp, cc) : FDTD2D_EINVAL;
}
int oracle_read_line(const oracle_sim *s, double *line, size_t count) {
  return s ? s->impl->read_line(line, count) : FDTD2D_EINVAL;
}
int oracle_read_rc(const oracle_sim *s, int32_t *spd, size_t count) {
  return s ? s->impl->read_rc(spd, count) : FDTD2D_EINVAL;
}
void oracle_destroy(oracle_sim *s) {
  if (s) {
    delete s->impl;
    delete s;
  }
}

}  // extern "C"
