// setup_internal.h - host-side owner of the arrays an fdtd2d_problem points to.
#ifndef FDTD2D_SETUP_INTERNAL_H
#define FDTD2D_SETUP_INTERNAL_H

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/fdtd2d.h"

namespace fdtd2d {

// Physical constants as the reference defines them (F/FDTD_cuda.cuh:20-23,
// F/FDTD_cuda.cu:503-504, F/kernel.cu:119).
constexpr double kPi = 3.14159265358979;
constexpr double kC = 2.99792458E8;
constexpr double kMu0 = 4.0 * kPi * 1.0E-7;
inline double eps0() { return 1.0 / (kC * kC * kMu0); }
constexpr float kCJ = 2.81794E-8f;  // e^2/m_e

// CPML grading parameters (F/FDTD_cuda.cu:1206-1214, F/FDTD_cuda.cuh:80).
struct CpmlParams {
  int npml = 11;
  int m = 3, ma = 1;
  double sigma_max_x = 0, sigma_max_y = 0;
  double alpha_max = 0.24;
  double kappa_max = 15.0;
};

// One axis' CPML tables: e-side [npml], h-side [npml-1], per side.
struct CpmlAxis {
  std::vector<float> be[2], ce[2], bh[2], ch[2], kappae[2], kappah[2];
};

struct Setup {
  fdtd2d_problem p{};
  std::vector<int32_t> material;
  std::vector<float> ca, cb, ne, density;
  CpmlAxis cx, cy;
  std::vector<float> den_ex, den_hx, den_ey, den_hy;
  std::vector<int32_t> src_i, src_j, det_i, det_j, comp_i, comp_j;
  std::vector<int32_t> rc_i, rc_imax;  // RC cutoff lines (scan column, maximum column)
  fdtd2d_processing proc{};            // [Processing], scaled to cells and steps
  std::vector<float> src_table;
  std::vector<float> source_profile;  // SourceP[nMax] envelope (SourceP.txt)
  std::map<std::string, double> params;
  std::string ini_log;  // iniLOG.txt text of the INI read (empty for other setups)
  std::string description;
  int ini_status = 0;
  int det_rows = 0, det_cols = 1;  // detector grid (det_qantI, det_qantK) for writers
  void bind();                     // point p at the vectors above
};

// Builds the CPML tables with the reference's mixed float/double arithmetic
// (F/FDTD_cuda.cu:1224-1459) and the den_* vectors (:1465-1578).
void build_cpml(const CpmlParams &cp, double dt, double dx, double dy, int nx, int ny, Setup &s);

// Material table of the reference (F/FDTD_cuda.cu:514-544, 1122-1200).
void build_reference_materials(double dt, float ne0, float nea, Setup &s);

}  // namespace fdtd2d

struct fdtd2d_setup {
  fdtd2d::Setup s;
};

#endif
