// snapshots.cpp - Ez snapshots of [2Dsave] (F/kernel.cu:1932-1933,
// 1983-1993, 2173-2302, 3041-3051): the cadence, and the text files the
// reference's viewers read (F/ImportScr.m, F/FieldView.py).
//
// With the reference's MATLAB_SAVE = 1 (F/Settings.cuh:120) a snapshot is,
// for each grid column i = 1, 1+GridSieve, ... < Imax-1, one line of Ez down
// that column (rows j = 1, 1+GridSieve, ... < Jmax-1) as "%12f " values; Ex,
// Hx, Hy and Hz.txt get one newline per such line and Ey.txt nothing;
// FieldSaveDat.txt holds "%d %d" = (lines per snapshot, snapshots).
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <string>
#include <thread>
#include <vector>

#include "errors.h"
#include "setup_internal.h"

using fdtd2d::set_error;

struct fdtd2d_snapshots {
  int nx = 0, ny = 0, gs = 1;
  int saved = 0;
  std::string dir;
  FILE *f[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};  // Ex Ey Ez Hx Hy Hz
  std::vector<std::string> rows;
};

namespace {

double param(const fdtd2d::Setup &S, const char *k, double def) {
  auto it = S.params.find(k);
  return it == S.params.end() ? def : it->second;
}

std::string in_dir(const std::string &dir, const char *name) {
  return dir.empty() ? std::string(name) : dir + "/" + name;
}

}  // namespace

extern "C" {

int fdtd2d_snapshot_steps(const fdtd2d_setup *setup, int nsteps_run, int32_t *steps, size_t capacity,
                          size_t *count) {
  if (!setup || !count || nsteps_run < 0) return set_error(FDTD2D_EINVAL, "bad argument");
  const fdtd2d::Setup &S = setup->s;
  const long from = static_cast<long>(param(S, "From", 0)), to = static_cast<long>(param(S, "To", 40000));
  const long sieve = static_cast<long>(param(S, "TimeSieve", 240));
  // the reference's counter starts at 1, counts every step inside [From, To]
  // and saves when it equals TimeSieve, then restarts at 1
  size_t c = 0;
  long app = 1;
  for (long n = 1; n <= nsteps_run; ++n) {
    if (n < from || n > to) continue;
    if (app == sieve) {
      if (steps && c < capacity) steps[c] = static_cast<int32_t>(n);
      ++c;
      app = 1;
    } else {
      ++app;
    }
  }
  *count = c;
  if (steps && c > capacity) return set_error(FDTD2D_EINVAL, "capacity below the snapshot count");
  return FDTD2D_OK;
}

int fdtd2d_snapshots_open(const fdtd2d_setup *setup, const char *dir, fdtd2d_snapshots **out) {
  if (!setup || !dir || !out) return set_error(FDTD2D_EINVAL, "null argument");
  const fdtd2d::Setup &S = setup->s;
  auto *w = new (std::nothrow) fdtd2d_snapshots();
  if (!w) return set_error(FDTD2D_ENOMEM, "out of host memory");
  w->nx = S.p.nx;
  w->ny = S.p.ny;
  w->gs = static_cast<int>(param(S, "GridSieve", 1));
  w->dir = dir;
  if (w->gs < 1) {
    delete w;
    return set_error(FDTD2D_EINVAL, "[2Dsave] GridSieve must be >= 1 (the reference's loops never end otherwise)");
  }
  const char *names[6] = {"Ex.txt", "Ey.txt", "Ez.txt", "Hx.txt", "Hy.txt", "Hz.txt"};
  for (int k = 0; k < 6; ++k) {
    w->f[k] = std::fopen(in_dir(w->dir, names[k]).c_str(), k < 2 ? "wb" : "w");
    if (!w->f[k]) {
      for (int q = 0; q < k; ++q) std::fclose(w->f[q]);
      delete w;
      return set_error(FDTD2D_EIO, std::string("cannot write ") + names[k] + " in " + dir);
    }
  }
  *out = w;
  return FDTD2D_OK;
}

int fdtd2d_snapshots_add(fdtd2d_snapshots *w, const float *ez, size_t count) {
  if (!w || !ez) return set_error(FDTD2D_EINVAL, "null argument");
  if (count != static_cast<size_t>(w->nx) * w->ny) return set_error(FDTD2D_EINVAL, "count must be nx*ny");
  std::vector<int> cols;
  for (int i = 1; i < w->nx - 1; i += w->gs) cols.push_back(i);
  const int nr = static_cast<int>(cols.size());
  w->rows.resize(nr);
  // one text line per column, formatted on host threads
  std::atomic<int> next{0};
  auto body = [&] {
    char buf[48];
    for (int r; (r = next.fetch_add(1)) < nr;) {
      std::string &s = w->rows[r];
      s.clear();
      const int i = cols[r];
      for (int j = 1; j < w->ny - 1; j += w->gs) {
        const int len = std::snprintf(buf, sizeof buf, "%12f ", static_cast<double>(ez[static_cast<size_t>(j) * w->nx + i]));
        s.append(buf, len);
      }
      s += '\n';
    }
  };
  const int nt = std::max(1, std::min<int>({16, nr, static_cast<int>(std::thread::hardware_concurrency())}));
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(body);
  body();
  for (auto &t : th) t.join();
  for (int r = 0; r < nr; ++r) {
    if (std::fwrite(w->rows[r].data(), 1, w->rows[r].size(), w->f[2]) != w->rows[r].size())
      return set_error(FDTD2D_EIO, "writing Ez.txt failed");
    for (int k : {0, 3, 4, 5}) std::fputc('\n', w->f[k]);
  }
  ++w->saved;
  return FDTD2D_OK;
}

int fdtd2d_snapshots_close(fdtd2d_snapshots *w, int *saved) {
  if (!w) return set_error(FDTD2D_EINVAL, "null argument");
  int rc = FDTD2D_OK;
  for (auto *f : w->f)
    if (f && std::fclose(f) != 0) rc = set_error(FDTD2D_EIO, "closing a snapshot file failed");
  int lines = 0;
  for (int i = 1; i < w->nx - 1; i += w->gs) ++lines;
  FILE *d = std::fopen(in_dir(w->dir, "FieldSaveDat.txt").c_str(), "w");
  if (!d) {
    rc = set_error(FDTD2D_EIO, "cannot write FieldSaveDat.txt in " + w->dir);
  } else {
    std::fprintf(d, "%d %d", lines, w->saved);
    std::fclose(d);
  }
  if (saved) *saved = w->saved;
  delete w;
  return rc;
}

}  // extern "C"
