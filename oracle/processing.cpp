// processing.cpp - CPU restatement of the reference's time-of-flight search
// (processing_core and its helpers, FDTD_2D_Full/processing.cu, abbreviated
// F/processing.cu).
//
// TEST INFRASTRUCTURE ONLY: the checker for the library's own implementation
// (fdtd2d_processing_tof). Written as the reference's structure reads, one
// function per reference function, with std::sort of every window in the
// median filter and vector copies where the reference copies; built with
// -ffp-contract=off.
//
// What pins it independently of this file (tests/test_processing.py):
//   moving_average  constant and linear series against their closed forms in
//                   the interior, the edge copies (including the value at
//                   w/2-1 the filter never writes) on hand-worked series, and
//                   against numpy's running float32 sum
//   medfilt         numpy's median over every window of random series,
//                   rank w/2 (upper median) for even windows
//   sec_diff        quadratics (constant second difference)
//   sd/cf/arc       hand-placed crossings on synthetic pulses, thresholds
//                   and the no-crossing return value
//   detector sim    uniform weights against a numpy box sum
// Where the reference reads outside its arrays (series shorter than a window
// or a delay + 1) this returns FDTD2D_EINVAL, as the library does.
#include <algorithm>
#include <cmath>
#include <vector>

#include "oracle.h"

namespace {

typedef std::vector<float> FloatVector;
typedef std::vector<FloatVector> FloatMatrix;

// F/processing.cu:170-189
void moving_average(FloatVector &src, FloatVector &dst, int win_size) {
  float buf = 0;
  int half_win = win_size / 2;
  int times = (int)src.size();
  for (int t = 0; t < win_size - 1; t++) buf += src[t];
  for (int t = win_size - 1; t < times; t++) {
    buf += src[t];
    dst[t - half_win] = buf / win_size;
    buf -= src[t - win_size + 1];
  }
  for (int t = half_win; t > 0; t--) {
    dst[t] = dst[half_win - 1];
    dst[times - t] = dst[times - half_win - 1];
  }
}

// F/processing.cu:191-213
void medfilt(FloatVector &src, FloatVector &dst, int win_size) {
  FloatVector buf, tmp;
  int half_win = win_size / 2;
  int times = (int)src.size();
  for (int t = 0; t < win_size - 1; t++) buf.push_back(src[t]);
  for (int t = win_size - 1; t < times; t++) {
    buf.push_back(src[t]);
    tmp = buf;
    std::sort(tmp.begin(), tmp.end());
    dst[t - half_win] = tmp[half_win];
    buf.erase(buf.begin());
  }
  for (int t = half_win; t > 0; t--) {
    dst[t] = dst[half_win - 1];
    dst[times - t] = dst[times - half_win - 1];
  }
}

// F/processing.cu:146-168
void filter(FloatVector &src, FloatVector &dst, int filter_type, int win_size, int iters) {
  FloatVector tmp = src;
  for (int i = 0; i < iters; i++) {
    if (filter_type == 1) {
      moving_average(tmp, dst, win_size);
      tmp = dst;
    } else if (filter_type == 2) {
      medfilt(tmp, dst, win_size);
      tmp = dst;
    }
  }
}

// F/processing.cu:293-301
float vmax(const FloatVector &v) {
  float res = v[0];
  for (size_t i = 1; i < v.size(); i++)
    if (v[i] > res) res = v[i];
  return res;
}

// F/processing.cu:286-291
void sec_diff(FloatVector &src, FloatVector &dst) {
  dst[0] = dst[dst.size() - 1] = 0;
  for (size_t i = 1; i + 1 < src.size(); i++) dst[i] = src[i - 1] - 2 * src[i] + src[i + 1];
}

// F/processing.cu:215-228
int sd_method(FloatVector &src, const fdtd2d_processing &A) {
  FloatVector tmp1 = src, tmp2 = tmp1;
  float min_level = A.tau_threshold * vmax(src);
  sec_diff(src, tmp1);
  filter(tmp1, tmp2, A.sdf_type, A.sdf_win, A.sdf_runs);
  for (size_t i = 0; i + 1 < src.size(); i++)
    if (src[i] > min_level)
      if (tmp2[i] * tmp2[i + 1] < 0) return (int)i;
  return (int)src.size() - 1;
}

// F/processing.cu:230-244 (cf) and :246-260 (arc): the same search
int cf_like(FloatVector &src, int tau, float k, float thr) {
  float min_level = thr * vmax(src), p1, p2;
  for (int t = 0; t < (int)src.size() - (tau + 1); t++) {
    if (src[t] > min_level) {
      p1 = src[t] + k * src[t + tau];
      p2 = src[t + 1] + k * src[t + tau + 1];
      if (p1 * p2 < 0) return t;
    }
  }
  return (int)src.size() - 1;
}

bool window_ok(int type, int w, int runs, int times) {
  return (type != 1 && type != 2) || runs < 1 || (w >= 1 && w <= times);
}

}  // namespace

extern "C" int oracle_processing_tof(const fdtd2d_processing *ap, const float *line, int times,
                                     const float *weights, int pass, float *sd, float *cf, float *arc,
                                     float *original, float *filtered) {
  if (!ap || !line) return FDTD2D_EINVAL;
  const fdtd2d_processing &A = *ap;
  const int sz = A.size_i;
  if (times < 1 || sz < 1 || (pass != 0 && pass != 1)) return FDTD2D_EINVAL;
  if (pass == 1 && (A.detector_size < 1 || A.detector_size >= sz)) return FDTD2D_EINVAL;
  if (!window_ok(A.filter_type, A.filter_win, A.filter_runs, times)) return FDTD2D_EINVAL;
  if (A.sec_der && !window_ok(A.sdf_type, A.sdf_win, A.sdf_runs, times)) return FDTD2D_EINVAL;
  if (A.cf && (A.cf_tau < 0 || A.cf_tau + 1 > times)) return FDTD2D_EINVAL;
  if (A.arc && (A.arc_tau < 0 || A.arc_tau + 1 > times)) return FDTD2D_EINVAL;

  // F/processing.cu:9-17: source[i][t] = |field[t][i]|
  FloatMatrix source(sz), filt;
  for (int i = 0; i < sz; i++) {
    source[i] = FloatVector(times);
    for (int t = 0; t < times; t++) source[i][t] = std::fabs(line[(size_t)t * sz + i]);
  }
  FloatMatrix *series = &source;
  FloatMatrix real_detector;
  int n = sz;
  if (pass == 0) {
    filt = source;
  } else {
    // F/processing.cu:64-100
    n = sz - A.detector_size;
    real_detector = FloatMatrix(n, FloatVector(times, 0));
    filt = real_detector;
    FloatVector det_weights(A.detector_size, 1.0 / A.detector_size);
    if (weights) det_weights.assign(weights, weights + A.detector_size);
    int half_size = A.detector_size / 2;
    for (int t = 0; t < times; t++) {
      for (int i = 0; i < n; i++) {
        float buf = 0;
        for (int j = -half_size; j < A.detector_size - half_size; j++)
          buf += source[i + half_size + j][t] * det_weights[j + half_size];
        real_detector[i][t] = buf / A.detector_size;
      }
    }
    series = &real_detector;
  }
  for (int i = 0; i < n; i++) {
    filter((*series)[i], filt[i], A.filter_type, A.filter_win, A.filter_runs);
    if (sd) sd[i] = A.sec_der ? (float)(A.begin_after + sd_method(filt[i], A)) : 0.f;
    if (cf) cf[i] = A.cf ? (float)(A.begin_after + cf_like(filt[i], A.cf_tau, A.cf_k, A.tau_threshold)) : 0.f;
    if (arc) arc[i] = A.arc ? (float)(A.begin_after + cf_like(filt[i], A.arc_tau, A.arc_k, A.tau_threshold)) : 0.f;
    for (int t = 0; t < times; t++) {
      if (original) original[(size_t)i * times + t] = (*series)[i][t];
      if (filtered) filtered[(size_t)i * times + t] = filt[i][t];
    }
  }
  return FDTD2D_OK;
}
