// processing.cpp - time-of-flight search over the processing line and its
// output files: processing_core of the reference (F/processing.cu:6-144),
// run by the host on the line the stepper captured.
//
// Every series is one coordinate of the line over the captured steps. The
// coordinates are independent and run on host threads. The arithmetic is the
// reference's, in its order and in float: the moving average keeps its
// running float sum, the median is the element of rank w/2 of the window (a
// sorted window updated by one insertion and one removal per sample instead
// of a sort per sample: selection does no arithmetic, so the value is the
// same), and every quirk of the reference filters is kept: the edge copies
// read dst[w/2 - 1], which the filter itself may not have written, and the
// second-derivative post-filter writes into a copy of the undifferentiated
// series.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fdtd2d.h"
#include "errors.h"

namespace {

using fdtd2d::set_error;

constexpr double kTimePrint = 1e12;  // TIME_PRINT_COEF, F/Settings.cuh:13

// dst[1..half] <- dst[half-1] and dst[T-half..T-1] <- dst[T-half-1], both
// ends in one loop from the inside out (F/processing.cu:184-188, 208-212).
void copy_edges(float *dst, int T, int half) {
  for (int t = half; t > 0; --t) {
    dst[t] = dst[half - 1];
    dst[T - t] = dst[T - half - 1];
  }
}

// moving_average, F/processing.cu:170-189. Needs 1 <= w <= T.
void moving_average(const float *src, float *dst, int T, int w) {
  const int half = w / 2;
  const float fw = static_cast<float>(w);
  float sum = 0.f;
  for (int t = 0; t < w - 1; ++t) sum += src[t];
  for (int t = w - 1; t < T; ++t) {
    sum += src[t];
    dst[t - half] = sum / fw;
    sum -= src[t - w + 1];
  }
  copy_edges(dst, T, half);
}

// medfilt, F/processing.cu:191-213: the element of rank w/2 of the last w
// samples. Needs 1 <= w <= T.
void median(const float *src, float *dst, int T, int w, std::vector<float> &win) {
  const int half = w / 2;
  win.clear();
  for (int t = 0; t < w - 1; ++t) win.insert(std::upper_bound(win.begin(), win.end(), src[t]), src[t]);
  for (int t = w - 1; t < T; ++t) {
    win.insert(std::upper_bound(win.begin(), win.end(), src[t]), src[t]);
    dst[t - half] = win[half];
    // drop the oldest sample: among the equal values, the one with its bits
    const float old = src[t - w + 1];
    auto r = std::equal_range(win.begin(), win.end(), old);
    auto it = r.first;
    for (auto q = r.first; q != r.second; ++q)
      if (std::memcmp(&*q, &old, 4) == 0) {
        it = q;
        break;
      }
    win.erase(it);
  }
  copy_edges(dst, T, half);
}

struct Work {
  std::vector<float> tmp, win;
};

// filter, F/processing.cu:146-168: `runs` passes of one filter, each reading
// the previous pass's output; dst keeps what it held where a pass writes
// nothing, and stays untouched for other filter types.
void filter(const float *src, float *dst, int T, int type, int w, int runs, Work &k) {
  if (type != 1 && type != 2) return;
  k.tmp.assign(src, src + T);
  for (int r = 0; r < runs; ++r) {
    if (type == 1)
      moving_average(k.tmp.data(), dst, T, w);
    else
      median(k.tmp.data(), dst, T, w, k.win);
    std::copy(dst, dst + T, k.tmp.begin());
  }
}

float max_of(const float *v, int T) {
  float m = v[0];
  for (int t = 1; t < T; ++t)
    if (v[t] > m) m = v[t];
  return m;
}

// sd_method, F/processing.cu:215-228.
int sd_method(const float *src, int T, const fdtd2d_processing &a, Work &k, std::vector<float> &d2,
              std::vector<float> &post) {
  const float level = a.tau_threshold * max_of(src, T);
  d2.assign(T, 0.f);
  for (int i = 1; i < T - 1; ++i) d2[i] = src[i - 1] - 2 * src[i] + src[i + 1];
  d2[0] = 0.f;
  d2[T - 1] = 0.f;
  post.assign(src, src + T);
  filter(d2.data(), post.data(), T, a.sdf_type, a.sdf_win, a.sdf_runs, k);
  for (int i = 0; i < T - 1; ++i)
    if (src[i] > level && post[i] * post[i + 1] < 0) return i;
  return T - 1;
}

// cf_method / arc_method, F/processing.cu:230-260 (the same search with its
// own delay and attenuation).
int fraction_method(const float *src, int T, int tau, float kk, float thr) {
  const float level = thr * max_of(src, T);
  for (int t = 0; t < T - (tau + 1); ++t) {
    if (src[t] > level) {
      const float p1 = src[t] + kk * src[t + tau];
      const float p2 = src[t + 1] + kk * src[t + tau + 1];
      if (p1 * p2 < 0) return t;
    }
  }
  return T - 1;
}

int check(const fdtd2d_processing &a, int times, int pass) {
  if (times < 1) return set_error(FDTD2D_EINVAL, "processing: no captured steps");
  if (a.size_i < 1) return set_error(FDTD2D_EINVAL, "processing: empty line");
  if (pass != 0 && pass != 1) return set_error(FDTD2D_EINVAL, "processing: pass must be 0 or 1");
  if (pass == 1 && (a.detector_size < 1 || a.detector_size >= a.size_i))
    return set_error(FDTD2D_EINVAL, "processing: detector simulation needs 1 <= DetSize < SizeI");
  auto win_ok = [&](int type, int w, int runs) { return (type != 1 && type != 2) || runs < 1 || (w >= 1 && w <= times); };
  if (!win_ok(a.filter_type, a.filter_win, a.filter_runs))
    return set_error(FDTD2D_EINVAL, "processing: FilterWin must lie in [1, captured steps] (the reference "
                                    "reads outside the series otherwise)");
  if (a.sec_der && !win_ok(a.sdf_type, a.sdf_win, a.sdf_runs))
    return set_error(FDTD2D_EINVAL, "processing: SDFWin must lie in [1, captured steps]");
  if (a.cf && (a.cf_tau < 0 || a.cf_tau + 1 > times))
    return set_error(FDTD2D_EINVAL, "processing: CFTau must lie in [0, captured steps - 1]");
  if (a.arc && (a.arc_tau < 0 || a.arc_tau + 1 > times))
    return set_error(FDTD2D_EINVAL, "processing: ARCTau must lie in [0, captured steps - 1]");
  return FDTD2D_OK;
}

// Runs f(i) for i in [0, n) on up to 16 host threads.
template <class F>
void parallel_for(int n, F f) {
  const int nt = std::max(1, std::min<int>({16, n, static_cast<int>(std::thread::hardware_concurrency())}));
  std::atomic<int> next{0};
  auto body = [&] {
    for (int i; (i = next.fetch_add(1)) < n;) f(i);
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(body);
  body();
  for (auto &t : th) t.join();
}

struct Result {
  int n = 0, T = 0;
  std::vector<float> original, filtered;  // [n][T]
  std::vector<float> sd, cf, arc;
};

// One pass of processing_core: the line itself (pass 0, F/processing.cu:9-42)
// or the simulated detector (pass 1, :64-119).
void run_pass(const fdtd2d_processing &a, const float *line, int T, const float *weights, int pass, Result &R) {
  const int sz = a.size_i;
  const int n = pass == 0 ? sz : sz - a.detector_size;
  R.n = n;
  R.T = T;
  R.original.assign(static_cast<size_t>(n) * T, 0.f);
  R.sd.assign(n, 0.f);
  R.cf.assign(n, 0.f);
  R.arc.assign(n, 0.f);
  // source[i][t] = |line[t][i]| (:9-16)
  std::vector<float> source;
  std::vector<float> &src = pass == 0 ? R.original : source;
  src.resize(static_cast<size_t>(sz) * T);
  parallel_for(sz, [&](int i) {
    float *s = &src[static_cast<size_t>(i) * T];
    for (int t = 0; t < T; ++t) s[t] = std::fabs(line[static_cast<size_t>(t) * sz + i]);
  });
  if (pass == 0) {
    R.filtered = R.original;  // filtered = source: the filters start from a copy
  } else {
    // real_detector[i][t] = (sum_j source[i+j][t] * w[j]) / DetSize, j in
    // order (:88-98); filtered starts as the zero matrix (:66-67)
    const int D = a.detector_size;
    parallel_for(n, [&](int i) {
      float *o = &R.original[static_cast<size_t>(i) * T];
      for (int t = 0; t < T; ++t) {
        float buf = 0;
        for (int j = 0; j < D; ++j) buf += src[static_cast<size_t>(i + j) * T + t] * weights[j];
        o[t] = buf / static_cast<float>(D);
      }
    });
    R.filtered.assign(static_cast<size_t>(n) * T, 0.f);
  }
  parallel_for(n, [&](int i) {
    Work k;
    std::vector<float> d2, post;
    const float *o = &R.original[static_cast<size_t>(i) * T];
    float *f = &R.filtered[static_cast<size_t>(i) * T];
    filter(o, f, T, a.filter_type, a.filter_win, a.filter_runs, k);
    if (a.sec_der) R.sd[i] = static_cast<float>(a.begin_after + sd_method(f, T, a, k, d2, post));
    if (a.cf) R.cf[i] = static_cast<float>(a.begin_after + fraction_method(f, T, a.cf_tau, a.cf_k, a.tau_threshold));
    if (a.arc)
      R.arc[i] = static_cast<float>(a.begin_after + fraction_method(f, T, a.arc_tau, a.arc_k, a.tau_threshold));
  });
}

std::string path_in(const std::string &dir, const char *name) {
  return dir.empty() ? std::string(name) : dir + "/" + name;
}

// "%d %lf %lf" lines of a TOF vector (F/processing.cu:263-269).
int write_tof(const std::string &dir, const char *name, const std::vector<float> &v, int sieve, int from, double dt) {
  FILE *f = std::fopen(path_in(dir, name).c_str(), "w");
  if (!f) return set_error(FDTD2D_EIO, std::string("cannot write ") + name + " in " + dir);
  for (size_t i = 0; i < v.size(); i += sieve)
    std::fprintf(f, "%d %lf %lf\n", from + static_cast<int>(i), static_cast<double>(v[i]),
                 (dt * static_cast<double>(v[i])) * kTimePrint);
  std::fclose(f);
  return FDTD2D_OK;
}

// "%f\t" rows of a [n][T] matrix, one row per kept step (F/processing.cu:271-284).
int write_matrix(const std::string &dir, const char *name, const std::vector<float> &m, int n, int T, int t_sieve,
                 int d_sieve) {
  FILE *f = std::fopen(path_in(dir, name).c_str(), "w");
  if (!f) return set_error(FDTD2D_EIO, std::string("cannot write ") + name + " in " + dir);
  if (n > 0) {
    std::string row;
    char buf[64];
    for (int t = 0; t < T; t += t_sieve) {
      row.clear();
      for (int i = 0; i < n; i += d_sieve) {
        std::snprintf(buf, sizeof buf, "%f\t", static_cast<double>(m[static_cast<size_t>(i) * T + t]));
        row += buf;
      }
      row += '\n';
      std::fputs(row.c_str(), f);
    }
  }
  std::fclose(f);
  return FDTD2D_OK;
}

// DetWeights.txt as the reference reads it (F/processing.cu:73-86): up to
// DetSize values with operator>>. A stream already at its end before all are
// read makes every weight 1; a value that does not parse reads as 0 and ends
// the loop; only whitespace left ends it too, keeping the rest at 1/DetSize,
// as does an absent file.
std::vector<float> load_weights(const std::string &dir, int D, bool from_file) {
  std::vector<float> w(D, static_cast<float>(1.0 / D));
  if (!from_file) return w;
  std::ifstream in(path_in(dir, "DetWeights.txt"));
  for (int i = 0; in && i < D; ++i) {
    if (!in.eof()) {
      in >> w[i];
    } else {
      w.assign(D, 1.f);
      break;
    }
  }
  return w;
}

}  // namespace

extern "C" int fdtd2d_processing_tof(const fdtd2d_processing *a, const float *line, int times, const float *weights,
                                     int pass, float *sd, float *cf, float *arc, float *original, float *filtered) {
  if (!a || !line) return set_error(FDTD2D_EINVAL, "null argument");
  if (int rc = check(*a, times, pass)) return rc;
  std::vector<float> w;
  if (pass == 1 && !weights) {
    w.assign(a->detector_size, static_cast<float>(1.0 / a->detector_size));
    weights = w.data();
  }
  Result R;
  run_pass(*a, line, times, weights, pass, R);
  if (sd) std::copy(R.sd.begin(), R.sd.end(), sd);
  if (cf) std::copy(R.cf.begin(), R.cf.end(), cf);
  if (arc) std::copy(R.arc.begin(), R.arc.end(), arc);
  if (original) std::copy(R.original.begin(), R.original.end(), original);
  if (filtered) std::copy(R.filtered.begin(), R.filtered.end(), filtered);
  return FDTD2D_OK;
}

extern "C" int fdtd2d_write_processing_outputs(const fdtd2d_processing *a, const char *dir, const float *line,
                                               int times) {
  if (!a || !dir || !line) return set_error(FDTD2D_EINVAL, "null argument");
  if (a->dim_sieve < 1 || a->time_sieve < 1) return set_error(FDTD2D_EINVAL, "processing: sieves must be >= 1");
  if (int rc = check(*a, times, a->sim_detect ? 1 : 0)) return rc;
  const std::string d(dir);
  int rc;
  {
    Result R;
    run_pass(*a, line, times, nullptr, 0, R);
    if (a->save_original && (rc = write_matrix(d, "original_field.txt", R.original, R.n, R.T, a->time_sieve,
                                               a->dim_sieve)))
      return rc;
    if (a->save_filtered && (rc = write_matrix(d, "filtered_field.txt", R.filtered, R.n, R.T, a->time_sieve,
                                               a->dim_sieve)))
      return rc;
    if (a->sec_der && (rc = write_tof(d, "SD.txt", R.sd, a->dim_sieve, a->pos_i, a->dt))) return rc;
    if (a->cf && (rc = write_tof(d, "CF.txt", R.cf, a->dim_sieve, a->pos_i, a->dt))) return rc;
    if (a->arc && (rc = write_tof(d, "ARC.txt", R.arc, a->dim_sieve, a->pos_i, a->dt))) return rc;
  }
  if (a->sim_detect) {
    const std::vector<float> w = load_weights(d, a->detector_size, a->load_weights);
    Result R;
    run_pass(*a, line, times, w.data(), 1, R);
    const int from = a->pos_i + a->detector_size / 2;
    if (a->save_detector && a->save_original &&
        (rc = write_matrix(d, "detector_original_field.txt", R.original, R.n, R.T, a->time_sieve, a->dim_sieve)))
      return rc;
    if (a->save_detector && a->save_filtered &&
        (rc = write_matrix(d, "detector_filtered_field.txt", R.filtered, R.n, R.T, a->time_sieve, a->dim_sieve)))
      return rc;
    if (a->sec_der && (rc = write_tof(d, "det_SD.txt", R.sd, a->dim_sieve, from, a->dt))) return rc;
    if (a->cf && (rc = write_tof(d, "det_CF.txt", R.cf, a->dim_sieve, from, a->dt))) return rc;
    if (a->arc && (rc = write_tof(d, "det_ARC.txt", R.arc, a->dim_sieve, from, a->dt))) return rc;
  }
  return FDTD2D_OK;
}
