// medium.cpp - plasma medium builders of the reference (see medium.h).
#include "medium.h"

#include <cmath>
#include <limits>

#include "../../include/fdtd2d.h"
#include "errors.h"
#include "ini.h"

namespace fdtd2d {

void read_medium_settings(const IniFile &r, MediumParams &m) {
  // F/FDTD_cuda.cu:87-127
  m.build = r.get_int("medium", "Build", 0);
  m.cenI = r.get_int("medium", "Icen", 275);
  m.cenJ = r.get_int("medium", "Jcen", 305);
  m.A = static_cast<float>(r.get_real("medium", "A", 1));
  m.B = static_cast<float>(r.get_real("medium", "B", 1));
  m.C = static_cast<float>(r.get_real("medium", "C", 1));
  m.P = r.get_int("medium", "P", 1);
  m.Q = r.get_int("medium", "Q", 1);
  m.N = r.get_int("medium", "N", 1);
  m.R = r.get_int("medium", "Radius", 50);
  m.alpha = static_cast<float>(r.get_real("medium", "Alpha", 2.0));
  m.beta = static_cast<float>(r.get_real("medium", "Beta", 1.0));
  m.ne0 = static_cast<float>(r.get_real("medium", "ne0", 1e19));
  m.neA = static_cast<float>(r.get_real("medium", "neA", 0));
  m.torusR = r.get_int("medium", "TorusBR", 175);
  m.fluc_build = r.get_int("mediumFluc", "FlucBuild", 0);
  m.fcenI = r.get_int("mediumFluc", "FlucIcen", 275);
  m.fcenJ = r.get_int("mediumFluc", "FlucJcen", 305);
  m.fA = static_cast<float>(r.get_real("mediumFluc", "FlucA", 1));
  m.fB = static_cast<float>(r.get_real("mediumFluc", "FlucB", 1));
  m.fC = static_cast<float>(r.get_real("mediumFluc", "FlucC", 1));
  m.fP = r.get_int("mediumFluc", "FlucP", 1);
  m.fQ = r.get_int("mediumFluc", "FlucQ", 1);
  m.fN = r.get_int("mediumFluc", "FlucN", 1);
  m.angI = static_cast<float>(r.get_real("mediumFluc", "FlucAngleI", 0));
  m.angJ = static_cast<float>(r.get_real("mediumFluc", "FlucAngleJ", 0));
  m.angK = static_cast<float>(r.get_real("mediumFluc", "FlucAngleK", 0));
  m.sign = r.get_int("mediumFluc", "FlucSign", 1);
  m.fR = r.get_int("mediumFluc", "FlucRadius", 10);
  m.falpha = static_cast<float>(r.get_real("mediumFluc", "FlucAlpha", 2.0));
  m.fbeta = static_cast<float>(r.get_real("mediumFluc", "FlucBeta", 1.0));
  m.fne0 = static_cast<float>(r.get_real("mediumFluc", "FlucNe0", 1e19));
}

void scale_medium_settings(MediumParams &m, double ds) {
  auto sc = [&](int v) { return static_cast<int>(v * (1 / ds)); };
  m.cenI = sc(m.cenI);
  m.cenJ = sc(m.cenJ);
  m.cenK = 1;
  m.R = sc(m.R);
  m.fcenI = sc(m.fcenI);
  m.fcenJ = sc(m.fcenJ);
  m.fcenK = 1;
  m.fR = sc(m.fR);
}

namespace {

// ID3 writes of the reference's yeeCube / yeeCubeFace('z') / addYeeCube('z')
// (F/FDTD_cuda.cu:2662-2681, 2683-2712, 2805-2849): a 2x2 block of Ez cells
// with corner (I, J + K) - K is the 2D grid's unit z index. Cells beyond
// the grid (the reference writes past the array there) are skipped.
struct Id3 {
  int nx, ny;
  std::vector<int32_t> &id;
  bool in(int i, int j) const { return i >= 0 && i < nx && j >= 0 && j < ny; }
  void set_block(int I, int J, int K, int m) {
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di)
        if (in(I + di, J + K + dj)) id[static_cast<size_t>(J + K + dj) * nx + I + di] = m;
  }
  void add_block(int I, int J, int K, int m) {
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di) {
        if (!in(I + di, J + K + dj)) continue;
        int32_t &v = id[static_cast<size_t>(J + K + dj) * nx + I + di];
        if (v == 0) v = 8;
        v = (v + m) > 8 ? v + m : 8;
      }
  }
};

// Truncation of the reference's double -> int assignments; NaN and values
// outside int (x86 cvttsd2si) give INT_MIN.
int to_int(double v) {
  if (!(v > -2147483648.0 && v < 2147483648.0)) return std::numeric_limits<int>::min();
  return static_cast<int>(v);
}

// buildSphericalMedium, equation type 2 (SPHERE_MEDIUM_EQUOTION, F/Settings.cuh:163):
// F/FDTD_cuda.cu:3114-3200. Only the Ez-face ('z') pass feeds the stepper.
void sphere(const MediumParams &m, Id3 &g) {
  const int ne0 = 100;  // M_COEF
  const float ra = static_cast<float>(m.R), alpha = m.alpha, beta = m.beta;
  const int k = 0;
  for (int i = 0; i < g.nx - 1; i += 2)
    for (int j = 0; j < g.ny - 1; j += 2) {
      const double q = (m.P * std::pow(static_cast<float>(i) - m.cenI, 2)) / std::pow(m.A, 2) +
                       (m.Q * std::pow(static_cast<float>(j) - m.cenJ, 2)) / std::pow(m.B, 2) +
                       (m.N * std::pow(static_cast<float>(k) - m.cenK, 2)) / std::pow(m.C, 2);
      const float chis = static_cast<float>(std::sqrt(q));
      const int mod = to_int(ne0 * std::pow((1 - std::pow(chis / ra, alpha)), beta) + 8);
      if (mod >= 8) g.set_block(i, j, k, mod);
    }
}

// buildSphericalMediumFluc, type 2: F/FDTD_cuda.cu:3202-3310 ('z' pass).
void fluctuation(const MediumParams &m, const std::vector<float> &pne, Id3 &g) {
  int ne0 = 0;
  for (int i = 8; i < static_cast<int>(pne.size()) && pne[i] < m.fne0; ++i) ne0 = i - 8;
  const float ax = static_cast<float>(m.angI * 3.14159 / 180);
  const float ay = static_cast<float>(m.angJ * 3.14159 / 180);
  const float az = static_cast<float>(m.angK * 3.14159 / 180);
  const int k = 0;
  const float ra = static_cast<float>(m.fR);
  for (int i = 0; i < g.nx - 1; i += 2)
    for (int j = 0; j < g.ny - 1; j += 2) {
      const int di = i - m.fcenI, dj = j - m.fcenJ, dk = k - m.fcenK;
      const int i_ = to_int(dk * std::sin(ay) + di * std::cos(ay) * std::cos(az) - dj * std::cos(ay) * std::sin(az));
      const int j_ = to_int(di * (std::cos(ax) * std::sin(az) + std::cos(az) * std::sin(ax) * std::sin(ay)) +
                            dj * (std::cos(ax) * std::cos(az) - std::sin(ax) * std::sin(ay) * std::sin(az)) -
                            dk * std::cos(ay) * std::sin(ax));
      const int k_ = to_int(di * (std::sin(ax) * std::sin(az) - std::cos(ax) * std::cos(az) * std::sin(ay)) +
                            dj * (std::cos(az) * std::sin(ax) + std::cos(ax) * std::sin(ay) * std::sin(az)) +
                            dk * std::cos(ax) * std::cos(ay));
      const double q = (m.fP * std::pow(static_cast<float>(i_), 2)) / std::pow(m.fA, 2) +
                       (m.fQ * std::pow(static_cast<float>(j_), 2)) / std::pow(m.fB, 2) +
                       (m.fN * std::pow(static_cast<float>(k_), 2)) / std::pow(m.fC, 2);
      const float chis = static_cast<float>(std::sqrt(q));
      const int mod = to_int(std::floor(ne0 * std::pow((1 - std::pow(chis / ra, m.falpha)), m.fbeta)));
      if (mod >= 0) g.add_block(i, j, k, m.sign * mod);
    }
}

// buildTorusMedium, type 2: F/FDTD_cuda.cu:3313-3378. The loop variable k is
// a float 1 passed to yeeCube's int K, so the blocks land one row down.
void torus(const MediumParams &m, Id3 &g) {
  const float kPiRef = 3.14f;  // M_PI, F/Settings.cuh:7
  const float R0 = static_cast<float>(m.torusR), Z0 = static_cast<float>(m.cenK);
  const int ne0 = 100;
  const float nea = 8, alpha = m.alpha, beta = m.beta, a = static_cast<float>(m.R);
  const float k = 1;
  for (int i = 0; i < g.nx - 1; ++i)
    for (int j = 0; j < g.ny - 1; ++j) {
      float p_x = static_cast<float>(i - m.cenI), p_y = static_cast<float>(j - m.cenJ), p_z = k - Z0;
      float phi;
      if (p_x > 0) {
        phi = std::atan(p_y / p_x);
      } else {
        const int sign = (p_y >= 0) - (p_y < 0);
        phi = p_x < 0 ? std::atan(p_y / p_x) + sign * kPiRef : sign * (kPiRef / 2);
      }
      p_x = p_x * std::cos(phi) + p_y * std::sin(phi) - R0;
      p_z /= k;
      const float r_ = std::sqrt(p_x * p_x + p_z * p_z);
      const int mod = to_int((ne0 - nea) * std::pow(1 - std::pow(r_ / a, alpha), beta) + nea);
      if (mod >= 8) g.set_block(i, j, static_cast<int>(k), mod);
    }
}

}  // namespace

int build_reference_medium(const MediumParams &m, int nx, int ny, const std::vector<float> &ne,
                           std::vector<int32_t> &id3) {
  Id3 g{nx, ny, id3};
  if (m.build == 1) sphere(m, g);
  else if (m.build == 2) torus(m, g);
  else if (m.build != 0) return set_error(FDTD2D_EINVAL, "Settings.ini: [medium] Build must be 0, 1 or 2");
  if (m.fluc_build == 2) fluctuation(m, ne, g);
  else if (m.fluc_build != 0)
    return set_error(FDTD2D_EUNSUPPORTED, "Settings.ini: [mediumFluc] FlucBuild other than 0 or 2 builds nothing "
                                          "in the reference; only 2 is implemented");
  // Ids past the material table (the reference would read beyond pne) are
  // clamped to its last entry.
  const int top = static_cast<int>(ne.size()) - 1;
  for (auto &v : id3)
    if (v > top) v = top;
  return FDTD2D_OK;
}

}  // namespace fdtd2d
