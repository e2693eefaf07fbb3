// writers.cpp - the reference's detector output files, written by the host
// driver from the traces a plan returns (formats of F/kernel.cu:2409-2493,
// 2653-2676 and F/FDTD_cuda.cu:3647-3650).
#include <cmath>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <string>
#include <vector>

#include "errors.h"
#include "setup_internal.h"

namespace {

constexpr double kDetMult = 1000;          // DET_MULT, F/Settings.cuh:72
constexpr double kTimePrint = 1e12;        // TIME_PRINT_COEF, F/Settings.cuh:13

FILE *open_in(const std::string &dir, const char *name, const char *mode) {
  std::string path = dir.empty() ? std::string(name) : dir + "/" + name;
  return std::fopen(path.c_str(), mode);
}

// detAll<suffix>.txt (t [us], Ex, Ey (0), Ez, Hx, Hy sums, Hz (0), all x 1000)
// and detEz/detHx/detHy<suffix>.txt (one value per element per line, a blank
// line per step). The main pass (F/kernel.cu:2440-2493) ends detAll lines with
// "\n " and gives detP.txt the line breaks; the source-to-detector pass
// (:1838-1888) ends them with "\n" and leaves detP_S2D.txt empty.
int write_components(const fdtd2d::Setup &S, const std::string &d, const char *suffix, const float *comp,
                     bool main_pass) {
  using fdtd2d::set_error;
  const fdtd2d_problem &p = S.p;
  const int nMax = p.nsteps, ncomp = p.ncomp;
  const double dt = p.dt;
  const size_t plane = static_cast<size_t>(nMax + 1) * ncomp;
  const std::string sfx(suffix);
  FILE *f = open_in(d, ("detAll" + sfx + ".txt").c_str(), "w");
  if (!f) return set_error(FDTD2D_EIO, "cannot write detAll" + sfx + ".txt in " + d);
  for (int i = 0; i < nMax; ++i) {
    double s3 = 0, s4 = 0, s5 = 0, z = 0;
    for (int k = 0; k < ncomp; ++k) {
      size_t q = static_cast<size_t>(i) * ncomp + k;
      s3 += comp[q];
      s4 += comp[plane + q];
      s5 += comp[2 * plane + q];
    }
    std::fprintf(f, "%lf\t%12lf\t%12lf\t%12lf\t%12lf\t%12lf\t%12lf\n%s", (dt * i) * 1000000, z * kDetMult,
                 z * kDetMult, s3 * kDetMult, s4 * kDetMult, s5 * kDetMult, z * kDetMult, main_pass ? " " : "");
  }
  std::fclose(f);
  const std::string names[4] = {"detEz" + sfx + ".txt", "detHx" + sfx + ".txt", "detHy" + sfx + ".txt",
                                "detP" + sfx + ".txt"};
  FILE *fs[4];
  for (int q = 0; q < 4; ++q) {
    fs[q] = open_in(d, names[q].c_str(), "w");
    if (!fs[q]) return set_error(FDTD2D_EIO, "cannot write " + names[q]);
  }
  const int nq = main_pass ? 4 : 3;
  const int rows = ncomp / (S.det_cols > 0 ? S.det_cols : 1);
  const int cols = S.det_cols > 0 ? S.det_cols : 1;
  for (int i = 0; i < nMax; ++i) {
    for (int j = 0; j < rows; ++j) {
      for (int k = 0; k < cols; ++k)
        for (int q = 0; q < 3; ++q)
          std::fprintf(fs[q], "%12lf\t", comp[q * plane + static_cast<size_t>(i) * ncomp + j * cols + k] * kDetMult);
      for (int q = 0; q < nq; ++q) std::fprintf(fs[q], "\n");
    }
    for (int q = 0; q < nq; ++q) std::fprintf(fs[q], "\n");
  }
  for (int q = 0; q < 4; ++q) std::fclose(fs[q]);
  return FDTD2D_OK;
}

// Files the reference opens and never writes in this build's configuration.
int touch(const std::string &d, const char *name) {
  FILE *f = open_in(d, name, "w");
  if (!f) return fdtd2d::set_error(FDTD2D_EIO, std::string("cannot write ") + name + " in " + d);
  std::fclose(f);
  return FDTD2D_OK;
}

}  // namespace

extern "C" int fdtd2d_write_outputs(const fdtd2d_setup *setup, const char *dir, const float *det,
                                    const float *comp, double seconds) {
  using fdtd2d::set_error;
  if (!setup || !dir) return set_error(FDTD2D_EINVAL, "null argument");
  const fdtd2d::Setup &S = setup->s;
  const fdtd2d_problem &p = S.p;
  const int nMax = p.nsteps;
  const int ndet = p.ndet, ncomp = p.ncomp;
  const double dt = p.dt;
  const std::string d(dir);

  // Detect.txt: t [ps], sum of Ez over the detector array x 1000, and a third
  // column that is uninitialised in the reference (written 0 here).
  if (det) {
    FILE *f = open_in(d, "Detect.txt", "w");
    if (!f) return set_error(FDTD2D_EIO, "cannot write Detect.txt in " + d);
    for (int i = 0; i < nMax; ++i) {
      double sum = 0, sum1 = 0;
      for (int k = 0; k < ndet; ++k) sum += det[static_cast<size_t>(i) * ndet + k];
      std::fprintf(f, "%lf\t%12lf\t%12lf\n ", (dt * i) * kTimePrint, sum * kDetMult, sum1 * kDetMult);
    }
    std::fclose(f);
  }
  if (comp) {
    if (int rc = write_components(S, d, "", comp, true)) return rc;
  }
  // SourceP.txt: the source envelope, one "%f" line per step index 0..nMax-1.
  {
    FILE *f = open_in(d, "SourceP.txt", "w");
    if (!f) return set_error(FDTD2D_EIO, "cannot write SourceP.txt");
    for (int k = 0; k < nMax; ++k)
      std::fprintf(f, "%f\n", k < static_cast<int>(S.source_profile.size()) ? S.source_profile[k] : 0.f);
    std::fclose(f);
  }
  // log.txt (append): grid, time step and wall time.
  {
    FILE *f = open_in(d, "log.txt", "a");
    if (!f) return set_error(FDTD2D_EIO, "cannot append log.txt");
    char tdate[64], ttime[64];
    std::time_t now = std::time(nullptr);
    std::strftime(tdate, sizeof tdate, "%m/%d/%y", std::localtime(&now));
    std::strftime(ttime, sizeof ttime, "%H:%M:%S", std::localtime(&now));
    const int Imax = p.nx, Jmax = p.ny, Kmax = 1;
    std::fprintf(f, "-------------------------------------------------------------------------\n");
    std::fprintf(f, "OS date:\t\t\t\t%s\n", tdate);
    std::fprintf(f, "OS time:\t\t\t\t%s\n", ttime);
    std::fprintf(f, "-------------------------------------------------------------------------\n");
    std::fprintf(f,
                 "NET PARAMETERS \n Ex(%d, %d, %d) \n Ey(%d, %d, %d) \n Ez(%d, %d, %d) \n Hx(%d, %d, %d) \n "
                 "Hy(%d, %d, %d) \n Hz(%d, %d, %d)\n\n(dx ,dy, dz) - dt: (%f mm, %f mm, %f mm) - %f ns\n",
                 Imax - 1, Jmax, Kmax - 1, Imax, Jmax - 1, Kmax - 1, Imax, Jmax, Kmax, Imax, Jmax - 1, Kmax,
                 Imax - 1, Jmax, Kmax, Imax - 1, Jmax - 1, Kmax - 1, p.dx * 1000, p.dy * 1000, p.dx * 1000,
                 dt * 1e9);
    std::fprintf(f, "\nSIMULATION PARAMETERS\n Total Simulation time: %d time steps or %e seconds\n", nMax,
                 nMax * dt);
    auto get = [&](const char *k) {
      auto it = S.params.find(k);
      return it == S.params.end() ? 0.0 : it->second;
    };
    std::fprintf(f,
                 "\nSAVE DATA PARAMETER\nSlises saved: %d\nData saved from time %d to time %d \nTime sieve "
                 "parameter: %d \nNet sieve parameter: %d\nData multiplier: %f\n",
                 static_cast<int>(get("SlicesSaved")), static_cast<int>(get("From")), static_cast<int>(get("To")),
                 static_cast<int>(get("TimeSieve")),
                 static_cast<int>(get("GridSieve")), 1.0);
    std::fprintf(f,
                 "\nSORCE AND DETECTOR\n Source point: (%d, %d, %d)\n Detector point(%d, %d, %d)\n Detector "
                 "size: %dx%d points",
                 static_cast<int>(get("Isource")), static_cast<int>(get("Jsource")), 1,
                 static_cast<int>(get("Idetect")), static_cast<int>(get("Jdetect")), 1, 10, 10);
    std::fprintf(f, "\n\nCalculation on GPU\n Calculation time: %f seconds\n\n", seconds);
    std::fclose(f);
  }
  // deb.txt, the material id (ID3) map DevFieldCalculation prints after the
  // run (F/kernel.cu:2683, 2737-2751): one line per column i in 1..Imax-2 of
  // "%d " for rows j in 1..Jmax-2, then two blank-line pairs; a problem with
  // no id map (density medium) has none to print. materials.txt is opened
  // beside it and left empty (its writes are commented out, :2685-2731).
  if (!S.material.empty()) {
    FILE *f = open_in(d, "deb.txt", "w");
    if (!f) return set_error(FDTD2D_EIO, "cannot write deb.txt in " + d);
    const int Imax = p.nx, Jmax = p.ny;
    std::string line;
    char num[16];
    for (int i = 1; i < Imax - 1; ++i) {
      line.clear();
      for (int j = 1; j < Jmax - 1; ++j) {
        std::snprintf(num, sizeof num, "%d ", S.material[static_cast<size_t>(j) * Imax + i]);
        line += num;
      }
      line += '\n';
      std::fwrite(line.data(), 1, line.size(), f);
    }
    std::fputs("\n\n\n\n", f);
    std::fclose(f);
  }
  // files DevFieldCalculation opens and never writes in this configuration
  // (F/kernel.cu:1948-1993, 2685): the debug and line-detector files, the
  // Ex/Ey/Hz component files, the difference file (written afterwards by the
  // source-to-detector regime only) and the composite-detector file
  for (const char *name : {"test.txt", "SDetect.txt", "LineDetect.txt", "LineDetect2.txt", "detEx.txt", "detEy.txt",
                           "detHz.txt", "DifDetEx.txt", "CompDetect.txt", "materials.txt"})
    if (int rc = touch(d, name)) return rc;
  return FDTD2D_OK;
}

// iniLOG.txt (append): the configuration echo of the reference's INI read
// (logINIread, F/FDTD_cuda.cu:178-303), captured when the setup read it.
extern "C" int fdtd2d_write_ini_log(const fdtd2d_setup *setup, const char *dir) {
  using fdtd2d::set_error;
  if (!setup || !dir) return set_error(FDTD2D_EINVAL, "null argument");
  const std::string &text = setup->s.ini_log;
  if (text.empty()) return FDTD2D_OK;  // the INI could not be opened (the reference logs nothing then)
  FILE *f = open_in(std::string(dir), "iniLOG.txt", "a");
  if (!f) return set_error(FDTD2D_EIO, std::string("cannot append iniLOG.txt in ") + dir);
  std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  return FDTD2D_OK;
}

// RC tracker post-processing of the reference (F/kernel.cu:3180-3270, called
// at :2628-2644 after the step loop): RClineFilterPostTime (centred moving
// average of width FilterWindow, float sums in the reference's order),
// CalculateRCfromPwrLines (RC_Line.txt: largest raw row, largest filtered
// row, that times ds) and CalculateTaufromPwrLines (TAU_Line.txt: 2 i for the
// first i >= 2000 where the filtered row stops rising, and that times dt in
// ps), written with the iostream default formatting the reference uses; and
// RcDet_data.txt, the line positions (F/FDTD_cuda.cu:3492-3536).
extern "C" int fdtd2d_write_rc_outputs(const fdtd2d_setup *setup, const char *dir, const int32_t *spd) {
  using fdtd2d::set_error;
  if (!setup || !dir) return set_error(FDTD2D_EINVAL, "null argument");
  const fdtd2d::Setup &S = setup->s;
  const fdtd2d_problem &p = S.p;
  const int nMax = p.nsteps, Q = p.nrc;
  if (Q > 0 && !spd) return set_error(FDTD2D_EINVAL, "null RC traces");
  auto get = [&](const char *k, double d) {
    auto it = S.params.find(k);
    return it == S.params.end() ? d : it->second;
  };
  const std::string d(dir);
  auto path = [&](const char *name) { return d.empty() ? std::string(name) : d + "/" + name; };
  {
    std::ofstream out(path("RcDet_data.txt"));
    if (!out) return set_error(FDTD2D_EIO, "cannot write RcDet_data.txt");
    const int q = static_cast<int>(get("RC_qantI", 0)), i0 = static_cast<int>(get("RC_Istart", 0)),
              di = static_cast<int>(get("RC_di", 0));
    for (int i = 0; i < q; ++i) out << i0 + di * i << " ";
    out << std::endl;
    out << 0 << " ";  // one K position, K = 0
  }
  if (Q == 0) return FDTD2D_OK;
  const int w = static_cast<int>(get("FilterWindow", 200));
  // RClineFilterPostTime
  std::vector<float> fl(static_cast<size_t>(nMax) * Q, 0.f);
  const float weight = static_cast<float>(1.0 / std::abs(w));
  for (int k = 0; k < Q; ++k) {
    int w_beg = w / 2, w_end = -w / 2;
    for (int i = 0; i < nMax; ++i) {
      float acc = 0;
      for (int j = w_beg; j > w_end; j--)
        if (j > 0 && j < nMax) acc += static_cast<float>(spd[static_cast<size_t>(j) * Q + k]) * weight;
      fl[static_cast<size_t>(i) * Q + k] = acc;
      w_beg++;
      w_end++;
    }
  }
  const double ds = get("ds", 1.0);
  {
    std::ofstream out(path("RC_Line.txt"));
    if (!out) return set_error(FDTD2D_EIO, "cannot write RC_Line.txt");
    for (int k = 0; k < Q; ++k) {
      int temp = 0;
      float ft = 0;
      for (int i = 0; i < nMax; ++i)
        if (spd[static_cast<size_t>(i) * Q + k] > temp) temp = spd[static_cast<size_t>(i) * Q + k];
      for (int i = 0; i < nMax; ++i)
        if (fl[static_cast<size_t>(i) * Q + k] > ft) ft = fl[static_cast<size_t>(i) * Q + k];
      out << temp << " " << ft << " " << ft * ds << std::endl;
    }
  }
  {
    std::ofstream out(path("TAU_Line.txt"));
    if (!out) return set_error(FDTD2D_EIO, "cannot write TAU_Line.txt");
    for (int k = 0; k < Q; ++k) {
      float ft = 0;
      int i = 2000;
      do {
        ft = static_cast<float>(i * 2);
        i++;
      } while (i < nMax && fl[static_cast<size_t>(i) * Q + k] - fl[static_cast<size_t>(i - 1) * Q + k] >= 0);
      out << ft << " " << ft * p.dt * kTimePrint << std::endl;
    }
  }
  return FDTD2D_OK;
}

extern "C" int fdtd2d_write_s2d_outputs(const fdtd2d_setup *setup, const char *dir, const float *det,
                                        const float *comp) {
  using fdtd2d::set_error;
  if (!setup || !dir || !det || !comp) return set_error(FDTD2D_EINVAL, "null argument");
  const fdtd2d::Setup &S = setup->s;
  const fdtd2d_problem &p = S.p;
  const std::string d(dir);
  // Detect_S2D.txt: t [us] and the detector sum x 1000 (F/kernel.cu:1827-1834)
  FILE *f = open_in(d, "Detect_S2D.txt", "w");
  if (!f) return set_error(FDTD2D_EIO, "cannot write Detect_S2D.txt in " + d);
  for (int i = 0; i < p.nsteps; ++i) {
    double sum = 0;
    for (int k = 0; k < p.ndet; ++k) sum += det[static_cast<size_t>(i) * p.ndet + k];
    std::fprintf(f, "%lf\t%12lf\n ", (p.dt * i) * 1000000, sum * kDetMult);
  }
  std::fclose(f);
  if (int rc = write_components(S, d, "_S2D", comp, false)) return rc;
  for (const char *name : {"CompDetect_S2D.txt", "detEx_S2D.txt", "detEy_S2D.txt", "detHz_S2D.txt"})
    if (int rc = touch(d, name)) return rc;
  return FDTD2D_OK;
}

extern "C" int fdtd2d_write_s2d_difference(const fdtd2d_setup *setup, const char *dir, const float *comp,
                                           const float *comp_free) {
  using fdtd2d::set_error;
  if (!setup || !dir || !comp || !comp_free) return set_error(FDTD2D_EINVAL, "null argument");
  const fdtd2d_problem &p = setup->s.p;
  const int nMax = p.nsteps, ncomp = p.ncomp;
  const size_t plane = static_cast<size_t>(nMax + 1) * ncomp;
  FILE *f = open_in(dir, "DifDetEx.txt", "w");
  if (!f) return set_error(FDTD2D_EIO, std::string("cannot write DifDetEx.txt in ") + dir);
  // F/kernel.cu:2497-2536: per element the float difference, per step the
  // double sums of the Ez, Hx, Hy differences and the
  // magnitude of E x H from float components (Ex, Ey, Hz sums are 0)
  for (int i = 0; i < nMax; ++i) {
    double s[7] = {0, 0, 0, 0, 0, 0, 0};  // s[1..6] = detSum .. detSum6
    for (int k = 0; k < ncomp; ++k) {
      const size_t q = static_cast<size_t>(i) * ncomp + k;
      for (int c = 0; c < 3; ++c) {
        const float diff = comp[c * plane + q] - comp_free[c * plane + q];
        s[3 + c] += diff;
      }
    }
    const float eh_i = static_cast<float>(s[2] * s[6] - s[3] * s[5]);
    const float eh_j = static_cast<float>(-(s[1] * s[6] - s[3] * s[4]));
    const float eh_k = static_cast<float>(s[1] * s[5] - s[2] * s[4]);
    const double mag = std::sqrt(std::pow(eh_i, 2) + std::pow(eh_j, 2) + std::pow(eh_k, 2));
    std::fprintf(f, "%lf\t%12lf\t%12lf\t%12lf\t%12lf\t%12lf\t%12lf\t%12lf\n ", (p.dt * i) * 1000000,
                 s[1] * kDetMult, s[2] * kDetMult, s[3] * kDetMult, s[4] * kDetMult, s[5] * kDetMult, s[6] * kDetMult,
                 mag * kDetMult);
  }
  std::fclose(f);
  return FDTD2D_OK;
}
