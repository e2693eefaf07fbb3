// setup.cpp - host side of the stepper: Settings.ini -> problem (the role of
// readINIparam / recuclParametersToMetric / setUp / initializeCPML /
// buildSourceProfile / calculate*Darrai in F/FDTD_cuda.cu) and the seeded
// synthetic workloads used for parity tests and benchmarks.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <new>
#include <random>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "errors.h"
#include "ini.h"
#include "geometry.h"
#include "medium.h"
#include "setup_internal.h"

namespace fdtd2d {

void Setup::bind() {
  p.material = material.empty() ? nullptr : material.data();
  p.nmaterials = static_cast<int32_t>(ca.size());
  p.ca = ca.empty() ? nullptr : ca.data();
  p.cb = cb.empty() ? nullptr : cb.data();
  p.ne = ne.empty() ? nullptr : ne.data();
  p.density = density.empty() ? nullptr : density.data();
  for (int s = 0; s < 2; ++s) {
    p.be_x[s] = cx.be[s].data();
    p.ce_x[s] = cx.ce[s].data();
    p.bh_x[s] = cx.bh[s].data();
    p.ch_x[s] = cx.ch[s].data();
    p.be_y[s] = cy.be[s].data();
    p.ce_y[s] = cy.ce[s].data();
    p.bh_y[s] = cy.bh[s].data();
    p.ch_y[s] = cy.ch[s].data();
  }
  p.den_ex = den_ex.data();
  p.den_hx = den_hx.data();
  p.den_ey = den_ey.data();
  p.den_hy = den_hy.data();
  p.nsrc = static_cast<int32_t>(src_i.size());
  p.src_i = src_i.data();
  p.src_j = src_j.data();
  p.src_table = src_table.data();
  p.ndet = static_cast<int32_t>(det_i.size());
  p.det_i = det_i.data();
  p.det_j = det_j.data();
  p.ncomp = static_cast<int32_t>(comp_i.size());
  p.comp_i = comp_i.data();
  p.comp_j = comp_j.data();
  p.nrc = static_cast<int32_t>(rc_i.size());
  p.rc_i = rc_i.empty() ? nullptr : rc_i.data();
  p.rc_imax = rc_imax.empty() ? nullptr : rc_imax.data();
}

// One side of one axis, F/FDTD_cuda.cu:1226-1261 (x, side 1) and its three
// copies. Types follow the reference: sigma/alpha/kappa are float arrays, so
// sig/kappa + alpha is a float sum before the double dt/eps0 scaling.
static void cpml_side(const CpmlParams &cp, double sig_max, double dt, std::vector<float> &be,
                      std::vector<float> &ce, std::vector<float> &bh, std::vector<float> &ch,
                      std::vector<float> &kappae, std::vector<float> &kappah) {
  const int P = cp.npml;
  const double e0 = eps0();
  be.assign(P, 0.f);
  ce.assign(P, 0.f);
  kappae.assign(P, 1.f);
  bh.assign(std::max(P - 1, 0), 0.f);
  ch.assign(std::max(P - 1, 0), 0.f);
  kappah.assign(std::max(P - 1, 0), 1.f);
  for (int i = 0; i < P; ++i) {
    float sige = static_cast<float>(sig_max * std::pow((P - 1 - i) / (P - 1.0), cp.m));
    float alphae = static_cast<float>(cp.alpha_max * std::pow(float(i) / (P - 1.0), cp.ma));
    float kap = static_cast<float>(1.0 + (cp.kappa_max - 1.0) * std::pow((P - 1 - i) / (P - 1.0), cp.m));
    float sk = sige / kap + alphae;
    be[i] = static_cast<float>(std::exp(-sk * dt / e0));
    if (sige == 0.0f && alphae == 0.0f && i == P - 1)
      ce[i] = 0.0f;
    else
      ce[i] = static_cast<float>(sige * (be[i] - 1.0) / (sige + kap * alphae) / kap);
    kappae[i] = kap;
  }
  for (int i = 0; i < P - 1; ++i) {
    float sigh = static_cast<float>(sig_max * std::pow((P - 1 - i - 0.5) / (P - 1.0), cp.m));
    float alphah = static_cast<float>(cp.alpha_max * std::pow((i + 1 - 0.5) / (P - 1.0), cp.ma));
    float kap = static_cast<float>(1.0 + (cp.kappa_max - 1.0) * std::pow((P - 1 - i - 0.5) / (P - 1.0), cp.m));
    float sk = sigh / kap + alphah;
    bh[i] = static_cast<float>(std::exp(-sk * dt / e0));
    ch[i] = static_cast<float>(sigh * (bh[i] - 1.0) / (sigh + kap * alphah) / kap);
    kappah[i] = kap;
  }
}

void build_cpml(const CpmlParams &cp, double dt, double dx, double dy, int nx, int ny, Setup &s) {
  const int P = cp.npml;
  for (int sd = 0; sd < 2; ++sd) {
    if (P > 0) {
      cpml_side(cp, cp.sigma_max_x, dt, s.cx.be[sd], s.cx.ce[sd], s.cx.bh[sd], s.cx.ch[sd],
                s.cx.kappae[sd], s.cx.kappah[sd]);
      cpml_side(cp, cp.sigma_max_y, dt, s.cy.be[sd], s.cy.ce[sd], s.cy.bh[sd], s.cy.ch[sd],
                s.cy.kappae[sd], s.cy.kappah[sd]);
    } else {
      for (auto *v : {&s.cx.be[sd], &s.cx.ce[sd], &s.cx.bh[sd], &s.cx.ch[sd], &s.cy.be[sd],
                      &s.cy.ce[sd], &s.cy.bh[sd], &s.cy.ch[sd]})
        v->clear();
    }
  }
  // Denominators, F/FDTD_cuda.cu:1465-1578: den = 1/(kappa*d) inside the
  // strips (mirror-indexed on the high side), 1/d elsewhere.
  s.den_hx.assign(nx - 1, 0.f);
  s.den_ex.assign(nx - 1, 0.f);
  s.den_hy.assign(ny - 1, 0.f);
  s.den_ey.assign(ny - 1, 0.f);
  auto fill = [&](std::vector<float> &den, int n, double d, const std::vector<float> *kap1,
                  const std::vector<float> *kap2, int lowlim, int startii) {
    int ii = startii;
    for (int i = 0; i < n - 1; ++i) {
      if (P > 0 && i < lowlim) {
        den[i] = static_cast<float>(1.0 / ((*kap1)[i] * d));
      } else if (P > 0 && i >= n - P) {
        den[i] = static_cast<float>(1.0 / ((*kap2)[ii] * d));
        ii = ii - 1;
      } else {
        den[i] = static_cast<float>(1.0 / d);
      }
    }
  };
  fill(s.den_hx, nx, dx, &s.cx.kappah[0], &s.cx.kappah[1], P - 1, P - 2);
  fill(s.den_hy, ny, dy, &s.cy.kappah[0], &s.cy.kappah[1], P - 1, P - 2);
  fill(s.den_ex, nx, dx, &s.cx.kappae[0], &s.cx.kappae[1], P, P - 1);
  fill(s.den_ey, ny, dy, &s.cy.kappae[0], &s.cy.kappae[1], P, P - 1);
}

void build_reference_materials(double dt, float ne0, float nea, Setup &s) {
  const int nm = 150;  // numMaterials, F/FDTD_cuda.cuh:170
  const double e0 = eps0();
  std::vector<double> eps(nm, e0), mu(nm, kMu0), sigma(nm, 0.0), ro(nm, 0.0);
  eps[2] = 4.0 * e0;  sigma[2] = 0.005;
  eps[3] = 8.0 * e0;  sigma[3] = 3.96E7;  // aluminium
  eps[4] = 9.5 * e0;  sigma[4] = 5.76E7;  // copper
  eps[5] = 9.0 * e0;  sigma[5] = 2e6;     // steel
  eps[6] = 2.1 * e0;  sigma[6] = 7.8e-4;  // teflon
  eps[7] = 81 * e0;   sigma[7] = 1e-2;    // water
  s.ca.assign(nm, 0.f);
  s.cb.assign(nm, 0.f);
  s.ne.assign(nm, 0.f);
  for (int i = 8; i < nm; ++i) s.ne[i] = static_cast<float>(((i - 8) * 0.01) * (ne0 - nea) + nea);
  float db0 = 0.f;
  for (int i = 0; i < nm; ++i) {
    s.ca[i] = static_cast<float>((1.0 - sigma[i] * dt / (2.0 * eps[i])) / (1.0 + sigma[i] * dt / (2.0 * eps[i])));
    s.cb[i] = static_cast<float>((dt / eps[i]) / (1.0 + sigma[i] * dt / (2.0 * eps[i])));
    float da = static_cast<float>((1.0 - ro[i] * dt / (2.0 * mu[i])) / (1.0 + ro[i] * dt / (2.0 * mu[i])));
    float db = static_cast<float>((dt / mu[i]) / (1.0 + ro[i] * dt / (2.0 * mu[i])));
    (void)da;  // == 1 for every material
    if (i == 0) db0 = db;
  }
  s.p.db = db0;
  s.p.pec_material = 1;
}

// ---------------------------------------------------------------- Settings.ini
namespace {

// Compile-time defaults of the reference, F/Settings.cuh.
struct RefSettings {
  int GPU_device = 0, S2D = 0;
  double ds = 1, dt_coef = 0.5;
  int Imax = 550, Jmax = 600;
  int Exp_time_key = 0, nMax = 10000;
  double Exp_time = 10;
  int srcI = 300, srcJ = 20, srcK = 1;
  int detI = 250, detJ = 30, detK = 1;
  int sPI = 40, sPK = 1, sAuto = 0, sDI = 5, sDK = 5;
  float angle = 0;
  int dPI = 40, dPK = 1, dAuto = 0, dDI = 5, dDK = 5;
  int chirp = 0;
  float freqA = 28, freqB = 30;
  int front = 2000, plato = 2000;
  double fronts_ns = 3, plato_ns = 3, amp = 1000;
  int fromT = 0, toT = 40000, timeSieve = 240, fieldSieve = 1;
  MediumParams med;
  int draw2d = 0, preset = 7, patL = 5, drawWall = 0, wallPos = 950;
  int rcI = 300, rcPI = 40, rcDI = 5;
  float rcLevel = 0.1f;
  int rcWindow = 200;
  fdtd2d_processing proc{};  // the reference's Processing global starts zeroed
};

void read_settings(const IniFile &r, RefSettings &s) {
  // F/FDTD_cuda.cu:35-172 (keys, sections and defaults as there)
  s.GPU_device = r.get_int("hardware", "GPU_device", 0);
  s.S2D = r.get_int("regimes", "SourceToDetect", 0);
  s.ds = r.get_real("grid", "ds", 1);
  s.dt_coef = r.get_real("grid", "dt_coef", 0.5);
  s.Imax = r.get_int("grid", "Imax", 550);
  s.Jmax = r.get_int("grid", "Jmax", 600);
  s.Exp_time_key = r.get_int("grid", "Use_Exp_time", 0) != 0;  // a bool in the reference
  s.nMax = r.get_int("grid", "Max_time", 10000);
  s.Exp_time = r.get_real("grid", "Exp_time", 10);
  s.srcI = r.get_int("grid", "Isource", 300);
  s.srcJ = r.get_int("grid", "Jsource", 20);
  s.detI = r.get_int("grid", "Idetect", 250);
  s.detJ = r.get_int("grid", "Jdetect", 30);
  s.sPI = r.get_int("AESA", "SIsize", 40);
  s.sAuto = r.get_int("AESA", "SAutoDelta", 0);
  s.sDI = r.get_int("AESA", "SIdelta", 5);
  s.angle = static_cast<float>(r.get_real("AESA", "SPAngle", 0));
  s.dPI = r.get_int("AESA", "DIsize", 40);
  s.dAuto = r.get_int("AESA", "SAutoDelta", 0);  // the reference reads SAutoDelta here too
  s.dDI = r.get_int("AESA", "DIdelta", 5);
  s.chirp = r.get_int("PulseParam", "Chirp", 0);
  s.freqA = static_cast<float>(r.get_real("PulseParam", "FreqA", 28));
  s.freqB = static_cast<float>(r.get_real("PulseParam", "FreqB", 30));
  s.front = r.get_int("PulseParam", "Fronts", 2000);
  s.plato = r.get_int("PulseParam", "Plato", 2000);
  s.fronts_ns = r.get_real("PulseParam", "Fronts_ns", 3);
  s.plato_ns = r.get_real("PulseParam", "Plato_ns", 3);
  s.amp = r.get_real("PulseParam", "Pulse_amp", 1000);
  s.fromT = r.get_int("2Dsave", "From", 0);
  s.toT = r.get_int("2Dsave", "To", 40000);
  s.timeSieve = r.get_int("2Dsave", "TimeSieve", 240);
  s.fieldSieve = r.get_int("2Dsave", "GridSieve", 1);
  read_medium_settings(r, s.med);
  s.draw2d = r.get_int("KameraConf", "Draw2D", 0);
  s.preset = r.get_int("KameraConf", "2Dpreset", 7);
  s.patL = r.get_int("KameraConf", "PatrubokL", 5);
  s.drawWall = r.get_int("KameraConf", "DrawWall", 0);
  s.wallPos = r.get_int("KameraConf", "WallPos", 950);
  s.rcI = r.get_int("RC_calc", "Iline", 300);
  s.rcPI = r.get_int("RC_calc", "LIsize", 40);
  s.rcDI = r.get_int("RC_calc", "LIdelta", 5);
  s.rcLevel = static_cast<float>(r.get_real("RC_calc", "Rlevel", 0.1));
  s.rcWindow = r.get_int("RC_calc", "FilterWindow", 200);
  // [Processing], F/FDTD_cuda.cu:144-172 (SDF* default to the prefilter's keys)
  fdtd2d_processing &a = s.proc;
  a.perform = r.get_bool("Processing", "Perform", false);
  a.sec_der = r.get_bool("Processing", "SecDer", false);
  a.cf = r.get_bool("Processing", "CF", false);
  a.arc = r.get_bool("Processing", "ARC", false);
  a.save_filtered = r.get_bool("Processing", "SaveFiltered", false);
  a.save_original = r.get_bool("Processing", "SaveOriginal", false);
  a.sim_detect = r.get_bool("Processing", "DetSim", false);
  a.save_detector = r.get_bool("Processing", "DetSave", false);
  a.load_weights = r.get_bool("Processing", "DetWLoad", false);
  auto gi = [&](const char *k, long d) { return static_cast<int32_t>(r.get_int("Processing", k, d)); };
  a.detector_size = gi("DetSize", 30);
  a.filter_type = gi("FilterType", 1);
  a.dim_sieve = gi("DimSieve", 1);
  a.time_sieve = gi("TimeSieve", 1);
  a.begin_after = gi("BeginAfter", 0);
  a.finish_at = gi("FinishAt", 0);
  a.pos_i = gi("PosI", 0);
  a.pos_j = gi("PosJ", 0);
  a.pos_k = gi("PosK", 0);
  a.size_i = gi("SizeI", 0);
  a.filter_win = gi("FilterWin", 300);
  a.filter_runs = gi("FilterRuns", 1);
  a.sdf_win = gi("SDFWin", a.filter_win);
  a.sdf_type = gi("SDFType", a.filter_type);
  a.sdf_runs = gi("SDFRuns", a.filter_runs);
  a.cf_tau = gi("CFTau", 300);
  a.arc_tau = gi("ARCTau", 300);
  a.tau_threshold = static_cast<float>(r.get_real("Processing", "TauThreshold", 0));
  a.arc_k = static_cast<float>(r.get_real("Processing", "ARCK", -0.5));
  a.cf_k = static_cast<float>(r.get_real("Processing", "CFK", -0.5));
}

// Element positions of calculate2Darrai / calculateDetect2Darrai
// (F/FDTD_cuda.cu:3380-3490) with the compile-time DS = 1 divisor, the
// disabled auto spacing (DI_AUTO = DD_AUTO = 0) and K = 1.
void aesa_array(int PI, int PK, int DI, int DK, int cenI, int cenK, std::vector<int> &arrI,
                std::vector<int> &arrK) {
  const int DS = 1;
  int sizeI = PI / DS, sizeK = PK / DS;
  int di = DI / DS, dk = DK / DS;
  if (di <= 0 || dk <= 0) {
    arrI.clear();
    arrK.clear();
    return;
  }
  int qI = static_cast<int>(std::floor(double(sizeI / di))) + 1;
  int qK = static_cast<int>(std::floor(double(sizeK / dk))) + 1;
  int Istart = static_cast<int>(std::floor(double(cenI - sizeI / 2)));
  int Kstart = static_cast<int>(std::floor(double(cenK - sizeK / 2)));
  arrI.resize(qI);
  arrK.resize(qK);
  for (int i = 0; i < qI; ++i) arrI[i] = Istart + di * i;
  for (int k = 0; k < qK; ++k) arrK[k] = Kstart + dk * k;
}

// The configuration echo logINIread appends to iniLOG.txt after reading the
// INI, before any unit conversion (F/FDTD_cuda.cu:178-303): the same keys in
// the same order, each value printed by an ostream as the reference's
// variable's type prints it (bool as 0/1, float and double in the default
// 6-significant-digit format), and the time and date of the read as its
// _strtime_s / _strdate_s give them (HH:MM:SS, MM/DD/YY).
std::string ini_log_text(const RefSettings &s) {
  char tm_s[32], dt_s[32];
  const std::time_t now = std::time(nullptr);
  std::strftime(tm_s, sizeof tm_s, "%H:%M:%S", std::localtime(&now));
  std::strftime(dt_s, sizeof dt_s, "%m/%d/%y", std::localtime(&now));
  const MediumParams &m = s.med;
  const fdtd2d_processing &a = s.proc;
  const bool exp_time = s.Exp_time_key != 0;
  std::ostringstream o;
  o << "Config loaded from 'Settings.ini' on time " << tm_s << " date " << dt_s << ":\nHardware:"
    << "\n GPU_device=" << s.GPU_device << "\nRegimes: "
    << "\n SourceToDetect=" << s.S2D << "\nGrid param: "
    << "\n Use_Exp_time=" << exp_time << "\n Imax=" << s.Imax << "\n Jmax=" << s.Jmax << "\n Max_time=" << s.nMax
    << "\n Exp_time=" << s.Exp_time << "\n ds=" << s.ds << "\n dt_coef=" << s.dt_coef << "\n Isource=" << s.srcI
    << "\n Jsource=" << s.srcJ << "\n Idetect=" << s.detI << "\n Jdetect=" << s.detJ
    << "\nAESA\n SIsize=" << s.sPI << "\n SAutoDelta=" << s.sAuto << "\n SIdelta=" << s.sDI
    << "\n SPAngle=" << s.angle << "\n DIsize=" << s.dPI << "\n DAutoDelta=" << s.dAuto << "\n DIdelta=" << s.dDI
    << "\nPulseParam:"
    << "\n Chirp=" << s.chirp << "\n FreqA=" << s.freqA << "\n FreqB=" << s.freqB << "\n Fronts=" << s.front
    << "\n Plato=" << s.plato << "\n Fronts_ns=" << s.fronts_ns << "\n Plato_ns=" << s.plato_ns
    << "\n Pulse_amp=" << s.amp << "\n2Dsave:"
    << "\n From=" << s.fromT << "\n To=" << s.toT << "\n TimeSieve=" << s.timeSieve
    << "\n GridSieve=" << s.fieldSieve << "\nmedium:"
    << "\n Build=" << m.build << "\n Icen=" << m.cenI << "\n Jcen=" << m.cenJ
    << "\n\n A=" << m.A << "\n B=" << m.B << "\n C=" << m.C << "\n P=" << m.P << "\n Q=" << m.Q << "\n N=" << m.N
    << "\n\n Radius=" << m.R << "\n Alpha=" << m.alpha << "\n Beta=" << m.beta << "\n ne0=" << m.ne0
    << "\n neA=" << m.neA << "\nmediumFluc:"
    << "\n FlucBuild=" << m.fluc_build << "\n FlucIcen=" << m.fcenI << "\n FlucJcen=" << m.fcenJ
    << "\n\n FlucA=" << m.fA << "\n FlucB=" << m.fB << "\n FlucC=" << m.fC << "\n FlucP=" << m.fP
    << "\n FlucQ=" << m.fQ << "\n FlucN=" << m.fN
    << "\n\n FlucAngleI=" << m.angI << "\n FlucAngleJ=" << m.angJ << "\n FlucAngleK=" << m.angK
    << "\n FlucSign=" << m.sign
    << "\n\n FlucRadius=" << m.fR << "\n FlucAlpha=" << m.falpha << "\n FlucBeta=" << m.fbeta
    << "\n FlucNe0=" << m.fne0
    << "\n TorusBR=" << m.torusR << "\nKameraConf:"
    << "\n Draw2D=" << s.draw2d << "\n 2Dpreset=" << s.preset << "\n PatrubokL=" << s.patL
    << "\n DrawWall=" << s.drawWall << "\n WallPos=" << s.wallPos
    << "\n\nRC_calc:"
    << "\n Iline=" << s.rcI << "\n LIsize=" << s.rcPI << "\n LIdelta=" << s.rcDI << "\n FilterWindow=" << s.rcWindow
    << "\n Rlevel=" << s.rcLevel
    << "\n\nProcessing:"
    << "\n Perform=" << a.perform << "\n SecDer=" << a.sec_der << "\n SDFType=" << a.sdf_type
    << "\n SDFWin=" << a.sdf_win << "\n SDFRuns=" << a.sdf_runs << "\n CF=" << a.cf << "\n CFTau=" << a.cf_tau
    << "\n CFK=" << a.cf_k << "\n ARC=" << a.arc << "\n ARCTau=" << a.arc_tau << "\n ARCK=" << a.arc_k
    << "\n TauThreshold=" << a.tau_threshold << "\n FilterType=" << a.filter_type << "\n FilterWin=" << a.filter_win
    << "\n FilterRuns=" << a.filter_runs << "\n SaveFiltered=" << a.save_filtered
    << "\n SaveOriginal=" << a.save_original << "\n PosI=" << a.pos_i << "\n PosJ=" << a.pos_j
    << "\n PosK=" << a.pos_k << "\n SizeI=" << a.size_i << "\n BeginAfter=" << a.begin_after
    << "\n FinishAt=" << a.finish_at << "\n TimeSieve=" << a.time_sieve << "\n DimSieve=" << a.dim_sieve
    << "\n SimDetect=" << a.sim_detect << "\n DetSave=" << a.save_detector << "\n DetSize=" << a.detector_size
    << "\n DetWLoad=" << a.load_weights
    << "\n-----------------------------------------------------------------------------\n";
  return o.str();
}

}  // namespace

int setup_from_ini(const std::string &path, Setup &S, bool objects) {
  IniFile r(path);
  S.ini_status = r.status();
  RefSettings st;
  // A file that cannot be opened leaves every compile-time default in place
  // (F/FDTD_cuda.cu:30-33 returns before any Get call); parse errors do not.
  if (r.status() >= 0) {
    read_settings(r, st);
    S.ini_log = ini_log_text(st);  // the reference logs a read that opened the file
  }
  // recuclParametersToMetric, F/FDTD_cuda.cu:305-382: int fields are scaled by
  // a double 1/ds and truncated on assignment.
  const double ds = st.ds;
  const double dx = ds * 1E-3, dy = ds * 1E-3, dz = ds * 1E-3;
  auto sc = [&](int v) { return static_cast<int>(v * (1 / ds)); };
  int Imax = sc(st.Imax), Jmax = sc(st.Jmax);
  int srcI = sc(st.srcI), srcJ = sc(st.srcJ);
  int detI = sc(st.detI), detJ = sc(st.detJ);
  scale_medium_settings(st.med, ds);
  int sPI = sc(st.sPI), sDI = sc(st.sDI);
  int dPI = sc(st.dPI), dDI = sc(st.dDI);
  const double dt = (0.99 / (kC * std::sqrt(1.0 / (dx * dx) + 1.0 / (dy * dy) + 1.0 / (dz * dz)))) * st.dt_coef;
  int nMax = st.nMax, front = st.front, plato = st.plato;
  if (st.Exp_time_key == 1) {
    nMax = static_cast<int>((st.Exp_time * 1e-9) / dt);
    front = static_cast<int>((st.fronts_ns * 1e-9) / dt);
    plato = static_cast<int>((st.plato_ns * 1e-9) / dt);
  }
  if (Imax < 2 * 11 + 4 || Jmax < 2 * 11 + 4 || nMax < 0)
    return set_error(FDTD2D_EINVAL, "Settings.ini: grid too small or negative step count");

  fdtd2d_problem &p = S.p;
  p.nx = Imax;
  p.ny = Jmax;
  p.npml = 11;  // PML_SIZE, F/Settings.cuh:23
  p.nsteps = nMax;
  p.dt = dt;
  p.dx = dx;
  p.dy = dy;
  p.dx_f = static_cast<float>(dx);
  p.dy_f = static_cast<float>(dy);
  p.cj = kCJ;
  p.jz_decay = 1.0f;
  p.jz_dt = dt;
  p.medium_kind = FDTD2D_MEDIUM_TABLE;

  // setUp, F/FDTD_cuda.cu:1122-1218
  build_reference_materials(dt, st.med.ne0, st.med.neA, S);
  CpmlParams cp;
  cp.npml = 11;
  const double eta = std::sqrt(kMu0 / (eps0() * 1.0));
  cp.sigma_max_x = 0.75 * (0.8 * (cp.m + 1) / (dx * eta));
  cp.sigma_max_y = 0.75 * (0.8 * (cp.m + 1) / (dy * eta));
  build_cpml(cp, dt, dx, dy, Imax, Jmax, S);

  // buildObject, F/FDTD_cuda.cu:2602-2659, in its order: with objects, the
  // chamber figure (Draw2D == 1) and then the plasma medium over it; the plate
  // wall (DrawWall == 1) last. initlimits (:384-441) builds without objects
  // for the first pass of the source-to-detector regime and never sets the
  // wall flag in that regime.
  S.material.assign(static_cast<size_t>(Imax) * Jmax, 0);
  if (objects) {
    if (st.draw2d == 1) draw_chamber_figure(st.preset, st.patL, srcI, Imax, Jmax, S.material);
    int mrc = build_reference_medium(st.med, Imax, Jmax, S.ne, S.material);
    if (mrc) return mrc;
  }
  if (st.S2D != 1 && st.drawWall == 1) build_plate_wall(st.wallPos, Imax, Jmax, S.material);

  // AESA source and detector arrays with the element offset +e and row +1.
  std::vector<int> aI, aK, dI, dK;
  aesa_array(sPI, st.sPK, sDI, st.sDK, srcI, st.srcK, aI, aK);
  aesa_array(dPI, st.dPK, dDI, st.dDK, detI, st.detK, dI, dK);
  const int qantI = static_cast<int>(aI.size()), qantK = static_cast<int>(aK.size());
  const int dqI = static_cast<int>(dI.size()), dqK = static_cast<int>(dK.size());
  for (int e = 0; e < qantI; ++e)
    for (int k = 0; k < qantK; ++k) {
      S.src_i.push_back(aI[e] + e);
      S.src_j.push_back(srcJ + aK[k]);
    }
  for (int d = 0; d < dqI; ++d)
    for (int k = 0; k < dqK; ++k) {
      S.det_i.push_back(dI[d] + d);
      S.det_j.push_back(detJ + dK[k]);
    }
  // Component detector: detector row, *source* array columns, det_qantI
  // elements (F/kernel.cu:2157 passes D_arrayI with det_qantI).
  for (int d = 0; d < std::min(dqI, qantI); ++d)
    for (int k = 0; k < dqK; ++k) {
      S.comp_i.push_back(aI[d] + d);
      S.comp_j.push_back(detJ + dK[k]);
    }
  S.det_rows = dqI;
  S.det_cols = dqK;

  // RC cutoff lines, calculateRCdetect2Darrai (F/FDTD_cuda.cu:3492-3536) on
  // the scaled [RC_calc] keys (int /= ds truncates, :356-361; LIsize of K is
  // 1 and its step 5, :137, 140, so one K position, K = 0). Dev_RCcalc
  // (F/kernel.cu:968-1044) takes the maximum in column RCdetArrayI[e] and scans
  // column RCdetArrayI[e] + e; the maximum search starts at row
  // SourceCen.J + 150 (:2170); the edge level starts at 10.0 (:1513). Lines
  // outside the grid are dropped (the reference would read out of bounds).
  {
    const int rcI = static_cast<int>(st.rcI / ds);
    const int sizeI = static_cast<int>(st.rcPI / ds), di = static_cast<int>(st.rcDI / ds);
    if (di > 0) {
      const int q = sizeI / di + 1;
      const int istart = rcI - sizeI / 2;
      for (int e = 0; e < q; ++e) {
        const int im = istart + di * e, is = im + e;
        if (im < 0 || im >= Imax || is < 0 || is >= Imax) continue;
        S.rc_imax.push_back(im);
        S.rc_i.push_back(is);
      }
    }
    p.rc_jcut = srcJ + 150;
    p.rc_level = st.rcLevel;
    p.rc_edge0 = 10.0f;
    S.params["RC_Istart"] = rcI - sizeI / 2;
    S.params["RC_di"] = di;
    S.params["RC_qantI"] = di > 0 ? sizeI / di + 1 : 0;
  }

  // [Processing] in cells and steps (F/FDTD_cuda.cu:364-377: int /= ds and
  // int *= 1e-9/dt truncate), and the line the stepper captures for it
  // (F/FDTD_cuda.cu:471-483, F/kernel.cu:2138-2145). A line outside the grid,
  // or off the single K plane, is not captured: the reference would read
  // outside its field or store nothing.
  {
    fdtd2d_processing a = st.proc;
    a.pos_i = static_cast<int32_t>(a.pos_i / ds);
    a.pos_j = static_cast<int32_t>(a.pos_j / ds);
    a.pos_k = static_cast<int32_t>(a.pos_k / ds);
    a.size_i = static_cast<int32_t>(a.size_i / ds);
    if (st.Exp_time_key == 1) {
      a.begin_after = static_cast<int32_t>(a.begin_after * (1e-9 / dt));
      a.finish_at = static_cast<int32_t>(a.finish_at * (1e-9 / dt));
    }
    a.times = a.perform ? std::max(0, std::min(a.finish_at, nMax) - a.begin_after) : 0;
    a.dt = dt;
    S.proc = a;
    const bool inside = a.size_i > 0 && a.pos_i >= 0 && static_cast<long long>(a.pos_i) + a.size_i <= Imax &&
                        a.pos_j >= 0 && a.pos_j < Jmax && a.pos_k == 0 && a.begin_after >= 0;
    if (a.perform && a.times > 0 && inside) {
      p.line_size = a.size_i;
      p.line_i0 = a.pos_i;
      p.line_j = a.pos_j;
      p.line_begin = a.begin_after;
      p.line_end = a.finish_at;
    }
  }

  // buildSourceProfile, F/FDTD_cuda.cu:3539-3655 (the uninitialised tail and
  // the read at n == nMax are defined as 0 here).
  const int platoW = plato, frontW = front / 2;
  std::vector<float> namp(std::max(front, 0));
  for (int i = 0; i < 2 * frontW; ++i) {
    double up = std::pow(double(i - frontW), 2.);
    double lc = double(frontW) / double(2.5);
    double lo = std::pow(lc, 2.);
    namp[i] = static_cast<float>(st.amp * (std::exp(-1.0 * up / lo)));
  }
  S.source_profile.assign(nMax + 1, 0.f);
  {
    int i, j;
    for (j = 0, i = 0; (i < frontW) && (j < nMax); i++, j++) S.source_profile[i] = namp[i];
    for (i = frontW; (i < frontW + platoW) && (j < nMax); i++, j++)
      S.source_profile[i] = static_cast<float>(st.amp);
    for (i = 0; (i < frontW) && (j < nMax); i++, j++)
      S.source_profile[frontW + platoW + i] = namp[frontW + i];
  }
  const float fCoef = st.freqA * 1E9f;
  const float fCoef2 = (st.chirp == 1 ? st.freqB : st.freqA) * 1E9f;
  const float fPulseW = static_cast<float>((platoW + 2 * frontW) * dt);
  // Dev_AESASourceCalc + Dev_calculateChirp (F/kernel.cu:621-659, 694-698):
  // float phase, double 2*3.14 and double cos, amplitude / element count.
  const int nsrc = qantI * qantK;
  S.src_table.assign(static_cast<size_t>(nMax + 1) * nsrc, 0.f);
  const float dtf = static_cast<float>(dt);
  const float pexp = 1.0f;
  for (int n = 1; n <= nMax; ++n) {
    const float t = static_cast<float>(n) * dtf;
    const float amp_n = S.source_profile[n];
    for (int e = 0; e < qantI; ++e) {
      const float phi = 90 + (st.angle * e);
      const float beta = (fCoef2 - fCoef) * std::pow(fPulseW, -pexp);
      const float ph = beta / (1 + pexp) * std::pow(t, 1 + pexp) + fCoef * t + phi / 360;
      const float c = static_cast<float>(std::cos(2 * 3.14 * ph));
      const float v = (amp_n / static_cast<float>(qantK * qantI)) * c;
      for (int k = 0; k < qantK; ++k) S.src_table[static_cast<size_t>(n) * nsrc + e * qantK + k] = v;
    }
  }

  auto &P = S.params;
  P["GPU_device"] = st.GPU_device;
  P["SourceToDetect"] = st.S2D;
  P["ds"] = ds;
  P["dt"] = dt;
  P["dx"] = dx;
  P["Imax"] = Imax;
  P["Jmax"] = Jmax;
  P["nMax"] = nMax;
  P["Isource"] = srcI;
  P["Jsource"] = srcJ;
  P["Idetect"] = detI;
  P["Jdetect"] = detJ;
  P["qantI"] = qantI;
  P["qantK"] = qantK;
  P["det_qantI"] = dqI;
  P["det_qantK"] = dqK;
  P["fCoef"] = fCoef;
  P["fCoef2"] = fCoef2;
  P["fPulseW"] = fPulseW;
  P["Front"] = front;
  P["Plato"] = plato;
  P["amp"] = st.amp;
  P["Angle"] = st.angle;
  P["Build"] = st.med.build;
  P["FlucBuild"] = st.med.fluc_build;
  P["From"] = st.fromT;
  P["To"] = st.toT;
  P["TimeSieve"] = st.timeSieve;
  P["GridSieve"] = st.fieldSieve;
  P["Draw2D"] = st.draw2d;
  P["2Dpreset"] = st.preset;
  P["PatrubokL"] = st.patL;
  P["DrawWall"] = st.drawWall;
  P["WallPos"] = st.wallPos;
  P["objects"] = objects ? 1 : 0;
  P["Iline"] = static_cast<int>(st.rcI / ds);
  P["LIsize"] = static_cast<int>(st.rcPI / ds);
  P["LIdelta"] = static_cast<int>(st.rcDI / ds);
  P["Rlevel"] = st.rcLevel;
  P["FilterWindow"] = st.rcWindow;
  P["npml"] = p.npml;
  S.description = "Settings.ini: " + path;
  S.bind();
  return FDTD2D_OK;
}

// ---------------------------------------------------------------- workloads
namespace {

// Standard normal deviates from mt19937_64 by Box-Muller on 53-bit uniforms,
// so the sequence does not depend on the C++ library's distribution code.
struct Normal {
  std::mt19937_64 rng;
  bool have = false;
  double spare = 0;
  explicit Normal(uint64_t seed) : rng(seed) {}
  double uniform() { return (static_cast<double>(rng() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double operator()() {
    if (have) {
      have = false;
      return spare;
    }
    double u1 = uniform(), u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(u1));
    spare = r * std::sin(2 * M_PI * u2);
    have = true;
    return r * std::cos(2 * M_PI * u2);
  }
};

template <class F>
void parallel_rows(int n, F &&f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 256) nt = 1;
  std::vector<std::thread> th;
  int chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    int a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=, &f]() {
      for (int r = a; r < b; ++r) f(r);
    });
  }
  for (auto &x : th) x.join();
}

}  // namespace

// Unit-variance Gaussian random field with correlation exp(-r^2/l^2): white
// N(0,1) noise convolved with a separable Gaussian of width s = l/2 cells,
// kernel truncated at 4s and normalised to unit output variance.
std::vector<float> gaussian_field(int nx, int ny, double corr_cells, uint64_t seed) {
  std::vector<float> w(static_cast<size_t>(nx) * ny);
  // one generator per row, seeded from (seed, row): rows fill in parallel and
  // the field does not depend on the thread count
  parallel_rows(ny, [&](int j) {
    Normal g(seed ^ (0x9E3779B97F4A7C15ull * static_cast<uint64_t>(j + 1)));
    float *row = &w[static_cast<size_t>(j) * nx];
    for (int i = 0; i < nx; ++i) row[i] = static_cast<float>(g());
  });
  const double s = corr_cells / 2.0;
  const int R = std::max(1, static_cast<int>(std::ceil(4 * s)));
  std::vector<double> k(2 * R + 1);
  double n2 = 0;
  for (int q = -R; q <= R; ++q) {
    k[q + R] = std::exp(-0.5 * (q / s) * (q / s));
    n2 += k[q + R] * k[q + R];
  }
  for (auto &v : k) v /= std::sqrt(n2);
  std::vector<float> t(w.size());
  parallel_rows(ny, [&](int j) {
    const float *src = &w[static_cast<size_t>(j) * nx];
    float *dst = &t[static_cast<size_t>(j) * nx];
    for (int i = 0; i < nx; ++i) {
      double acc = 0;
      for (int q = -R; q <= R; ++q) {
        int ii = i + q;
        if (ii < 0) ii = -ii - 1;
        if (ii >= nx) ii = 2 * nx - ii - 1;
        ii = std::min(std::max(ii, 0), nx - 1);
        acc += k[q + R] * src[ii];
      }
      dst[i] = static_cast<float>(acc);
    }
  });
  parallel_rows(ny, [&](int j) {
    // accumulate whole rows, tap by tap (the same sum order per cell as a
    // per-cell loop over the taps, but row-contiguous)
    std::vector<double> acc(nx, 0.0);
    for (int q = -R; q <= R; ++q) {
      int jj = j + q;
      if (jj < 0) jj = -jj - 1;
      if (jj >= ny) jj = 2 * ny - jj - 1;
      jj = std::min(std::max(jj, 0), ny - 1);
      const float *src = &t[static_cast<size_t>(jj) * nx];
      const double kq = k[q + R];
      for (int i = 0; i < nx; ++i) acc[i] += kq * src[i];
    }
    float *dst = &w[static_cast<size_t>(j) * nx];
    for (int i = 0; i < nx; ++i) dst[i] = static_cast<float>(acc[i]);
  });
  return w;
}

namespace {

double critical_density(double f_hz) {
  const double w = 2 * kPi * f_hz;
  return eps0() * w * w / static_cast<double>(kCJ);
}

// Common skeleton of the synthetic workloads: 2D Courant dt x 0.99, CPML with
// m = 3, sigma_max = 0.8(m+1)/(eta0 dx), kappa_max = 1, alpha_max = 0.05
// graded as the reference grades alpha (ma = 1), DENSITY medium.
// dt = 0.99 of the 2D Courant limit, or, for a plasma up to density ne_max,
// 0.99 of the Yee-Drude limit (c dt)^2 (1/dx^2 + 1/dy^2) + (wp dt / 2)^2 <= 1
// (the explicit Jz/Ez leapfrog adds the plasma frequency to the lattice
// dispersion; at 0.99 of the plain Courant limit, shots above about 1.6 times
// the critical density diverge).
void workload_base(Setup &S, int nx, int ny, double dx, int npml, int nsteps, double nu, double ne_max = 0.0) {
  fdtd2d_problem &p = S.p;
  p.nx = nx;
  p.ny = ny;
  p.npml = npml;
  p.nsteps = nsteps;
  p.dx = dx;
  p.dy = dx;
  p.dx_f = static_cast<float>(dx);
  p.dy_f = static_cast<float>(dx);
  const double wp2 = ne_max * kCJ / eps0();
  p.dt = 0.99 / std::sqrt(kC * kC * (1.0 / (dx * dx) + 1.0 / (dx * dx)) + wp2 / 4.0);
  const double dt = p.dt;
  p.db = static_cast<float>(dt / kMu0);
  p.medium_kind = FDTD2D_MEDIUM_DENSITY;
  p.ca0 = 1.0f;
  p.cb0 = static_cast<float>(dt / eps0());
  p.pec_material = -1;
  p.cj = kCJ;
  p.jz_decay = static_cast<float>((1.0 - nu * dt / 2) / (1.0 + nu * dt / 2));
  p.jz_dt = dt / (1.0 + nu * dt / 2);
  CpmlParams cp;
  cp.npml = npml;
  const double eta0 = std::sqrt(kMu0 / eps0());
  cp.sigma_max_x = cp.sigma_max_y = 0.8 * (cp.m + 1) / (eta0 * dx);
  cp.alpha_max = 0.05;
  cp.kappa_max = 1.0;
  build_cpml(cp, dt, dx, dx, nx, ny, S);
  S.density.assign(static_cast<size_t>(nx) * ny, 0.f);
  S.params["dt"] = dt;
  S.params["dx"] = dx;
  S.params["nu"] = nu;
  S.params["npml"] = npml;
  S.params["Imax"] = nx;
  S.params["Jmax"] = ny;
  S.params["nMax"] = nsteps;
}

// Gaussian-modulated sinusoid per element with a steering delay:
// Ez += amp * exp(-((t-t0-d_e)/tau)^2) * sin(2 pi f (t-t0-d_e)), t = n dt,
// computed in double and stored as the subtracted value.
void gaussian_pulse_table(Setup &S, double f, double tau, double t0, double amp,
                          const std::vector<double> &delay) {
  const int nsrc = static_cast<int>(S.src_i.size());
  const int nsteps = S.p.nsteps;
  S.src_table.assign(static_cast<size_t>(nsteps + 1) * nsrc, 0.f);
  S.source_profile.assign(nsteps + 1, 0.f);
  for (int n = 1; n <= nsteps; ++n) {
    const double t = n * S.p.dt;
    for (int e = 0; e < nsrc; ++e) {
      const double u = t - t0 - delay[e];
      const double v = amp * std::exp(-(u / tau) * (u / tau)) * std::sin(2 * kPi * f * u);
      S.src_table[static_cast<size_t>(n) * nsrc + e] = static_cast<float>(-v);
    }
    const double u = t - t0;
    S.source_profile[n] = static_cast<float>(amp * std::exp(-(u / tau) * (u / tau)));
  }
  S.params["f0"] = f;
  S.params["tau"] = tau;
  S.params["t0"] = t0;
}

// Plasma density profiles along x, uniform or fluctuating in y.
void plasma_profile(Setup &S, int kind, double peak, double x0, double x1, double fluct,
                    double corr, uint64_t seed) {
  const int nx = S.p.nx, ny = S.p.ny;
  std::vector<double> prof(nx, 0.0);
  for (int i = 0; i < nx; ++i) {
    double v = 0;
    if (kind == 0) {  // linear ramp 0 at x0 -> peak at x1, held beyond
      if (i > x0) v = peak * std::min(1.0, (i - x0) / (x1 - x0));
    } else {  // parabola peaking at x1 with half-width (x1 - x0)
      double z = (i - x1) / (x1 - x0);
      v = peak * std::max(0.0, 1.0 - z * z);
    }
    prof[i] = v;
  }
  std::vector<float> g;
  if (fluct > 0) g = gaussian_field(nx, ny, corr, seed);
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      size_t c = static_cast<size_t>(j) * nx + i;
      double v = prof[i];
      if (fluct > 0) v *= (1.0 + fluct * g[c]);
      S.density[c] = static_cast<float>(std::max(0.0, v));
    }
}

uint64_t mix(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t z = seed * 0x9E3779B97F4A7C15ull + a * 0xBF58476D1CE4E5B9ull + b * 0x94D049BB133111EBull;
  z ^= z >> 31;
  z *= 0xD6E8FEB86659FD93ull;
  z ^= z >> 29;
  return z;
}

}  // namespace

int setup_workload(int workload, uint64_t seed, int shot, int steps_override, Setup &S) {
  const double c = kC;
  switch (workload) {
    case 1: {  // 400x400 plain method
      int nsteps = steps_override > 0 ? steps_override : 2000;
      workload_base(S, 400, 400, 1e-3, 12, nsteps, 1e9);
      const double f = 10e9, nc = critical_density(f);
      plasma_profile(S, 0, 2 * nc, 200, 388, 0, 0, 0);
      S.src_i = {60};
      S.src_j = {200};
      gaussian_pulse_table(S, f, 0.3e-9, 1.2e-9, 1.0, {0.0});
      S.det_i = {60, 60, 60, 60};
      S.det_j = {170, 190, 210, 230};
      S.comp_i = S.det_i;
      S.comp_j = S.det_j;
      S.params["nc"] = nc;
      S.description = "400x400 plain method, linear plasma ramp, 1 source, 4 detectors";
      break;
    }
    case 2: {  // 8192^2 single grid
      int nsteps = steps_override > 0 ? steps_override : 20000;
      const int N = 8192;
      const double dx = 0.5e-3;
      workload_base(S, N, N, dx, 16, nsteps, 5e8);
      const double f = 20e9, nc = critical_density(f);
      plasma_profile(S, 1, 1.5 * nc, 4000, 6000, 0.02, 20, mix(seed, 2, 0));
      const int ne_ = 16, pitch = 8, j0 = N / 2 - ne_ * pitch / 2;
      std::vector<double> delay(ne_);
      const double steer = 15.0 * kPi / 180.0;
      for (int e = 0; e < ne_; ++e) {
        S.src_i.push_back(200);
        S.src_j.push_back(j0 + pitch * e);
        delay[e] = e * pitch * dx * std::sin(steer) / c;
      }
      gaussian_pulse_table(S, f, 3.0 / f, 4 * 3.0 / f, 1.0, delay);
      for (int d = 0; d < 64; ++d) {
        S.det_i.push_back(200);
        S.det_j.push_back(j0 + 2 * d);
      }
      S.comp_i = {200};
      S.comp_j = {N / 2};
      S.params["nc"] = nc;
      S.description = "8192x8192 single grid, 16-element 15 deg steered array, parabolic plasma";
      break;
    }
    case 3:    // one shot of the 256-shot uneven batch
    case 4:    // one frequency of a served sweep
    case 5: {  // one shot of the mixed-size stream
      std::mt19937_64 rng(mix(seed, workload, workload == 4 ? shot / 64 : shot));
      auto U = [&](double a, double b) {
        return a + (b - a) * (static_cast<double>(rng() >> 11) * (1.0 / 9007199254740992.0));
      };
      auto UI = [&](int a, int b) { return a + static_cast<int>(rng() % static_cast<uint64_t>(b - a + 1)); };
      int nx, ny, npml = 12, nsteps, nel, profile;
      double f, dx, peak, fluct, steer, nu = 1e9;
      if (workload == 3) {
        nx = 256 + 128 * UI(0, 14);
        ny = 256 + 128 * UI(0, 14);
        f = U(8e9, 40e9);
        dx = c / f / 20;
        nel = UI(1, 16);
        steer = U(-20, 20);
        profile = UI(0, 1);
        peak = U(0.5, 3.0) * critical_density(f);
        fluct = U(0, 0.05);
        nsteps = 4000;
      } else if (workload == 4) {
        // request-level draws (shared grid and plasma), then this frequency
        int n = 512 + 128 * UI(0, 12);
        nx = ny = n;
        int nfreq = UI(8, 32);
        dx = c / 60e9 / 20;
        profile = 1;
        peak = U(0.5, 3.0) * critical_density(10e9);
        fluct = 0.0;
        nel = 8;
        steer = 0;
        int k = shot % 64;
        if (k >= nfreq) return set_error(FDTD2D_EINVAL, "sweep frequency index beyond the request's count");
        for (int q = 0; q < k; ++q) (void)rng();
        f = U(10e9, 60e9);
        nsteps = 3000;
        S.params["nfreq"] = nfreq;
      } else {
        bool big = U(0, 1) < 0.1;
        if (big) {
          nx = ny = 16384;
          npml = 16;
          nel = 32;
          nsteps = 30000;
          fluct = 0.03;
        } else {
          nx = ny = 1024;
          nel = UI(1, 16);
          nsteps = 4000;
          fluct = U(0, 0.05);
        }
        f = 20e9;
        dx = c / f / 20;
        profile = 1;
        peak = 1.5 * critical_density(f);
        steer = U(-20, 20);
        S.params["big"] = big ? 1 : 0;
      }
      if (steps_override > 0) nsteps = steps_override;
      if (steps_override < 0) {  // the draws only: no arrays are built
        S.params["Imax"] = nx;
        S.params["Jmax"] = ny;
        S.params["nMax"] = nsteps;
        S.params["f0"] = f;
        S.params["nel"] = nel;
        S.params["workload"] = workload;
        S.description = "draws of a synthetic shot (workload " + std::to_string(workload) + ")";
        return FDTD2D_OK;
      }
      // densest cell: the profile peak times the fluctuation's largest factor
      // (a unit-variance field stays within 6 sigma over these grid sizes)
      workload_base(S, nx, ny, dx, npml, nsteps, nu, peak * (1.0 + 6.0 * fluct));
      const double x0 = 0.25 * nx, x1 = profile == 0 ? nx - npml - 1 : 0.75 * nx;
      plasma_profile(S, profile, peak, x0, x1, fluct, 20, mix(seed, 100 + workload, shot));
      const int pitch = std::max(2, static_cast<int>(std::lround(0.5 * c / f / dx)));
      const int span = pitch * (nel - 1);
      int jstart = ny / 2 - span / 2;
      if (jstart < npml + 2 || jstart + span > ny - npml - 3) {
        const int p2 = std::max(1, (ny - 2 * npml - 6) / std::max(1, nel - 1));
        jstart = npml + 3;
        (void)p2;
      }
      const int isrc = npml + 8;
      std::vector<double> delay(nel);
      const double st = steer * kPi / 180.0;
      for (int e = 0; e < nel; ++e) {
        int j = jstart + pitch * e;
        j = std::min(std::max(j, npml + 2), ny - npml - 3);
        S.src_i.push_back(isrc);
        S.src_j.push_back(j);
        delay[e] = (st >= 0 ? e : (nel - 1 - e)) * pitch * dx * std::fabs(std::sin(st)) / c;
      }
      gaussian_pulse_table(S, f, 2.0 / f, 8.0 / f, 1.0, delay);
      for (int d = 0; d < 4; ++d) {
        S.det_i.push_back(isrc + 2);
        S.det_j.push_back(std::min(ny - npml - 3, std::max(npml + 2, ny / 2 - 6 + 4 * d)));
      }
      S.params["f0"] = f;
      S.params["nel"] = nel;
      S.params["peak"] = peak;
      S.params["fluct"] = fluct;
      S.params["profile"] = profile;
      S.params["steer"] = steer;
      S.description = "synthetic shot (workload " + std::to_string(workload) + ")";
      break;
    }
    default:
      return set_error(FDTD2D_EINVAL, "unknown workload");
  }
  S.params["workload"] = workload;
  S.bind();
  return FDTD2D_OK;
}

int setup_box(int nx, int ny, int npml, int nsteps, double ds, Setup &S) {
  if (nx < 2 * npml + 4 || ny < 2 * npml + 4 || npml == 1 || npml < 0 || nsteps < 0 || ds <= 0)
    return set_error(FDTD2D_EINVAL, "setup_box: bad sizes");
  workload_base(S, nx, ny, ds, npml, nsteps, 0.0);
  S.src_table.clear();
  S.source_profile.assign(nsteps + 1, 0.f);
  S.description = "box";
  S.bind();
  return FDTD2D_OK;
}

}  // namespace fdtd2d

// ---------------------------------------------------------------- C ABI
using fdtd2d::set_error;

extern "C" {

int fdtd2d_setup_from_ini(const char *ini_path, fdtd2d_setup **out) {
  return fdtd2d_setup_from_ini_objects(ini_path, 1, out);
}

int fdtd2d_setup_from_ini_objects(const char *ini_path, int objects, fdtd2d_setup **out) {
  if (!ini_path || !out) return set_error(FDTD2D_EINVAL, "null argument");
  auto *s = new (std::nothrow) fdtd2d_setup();
  if (!s) return set_error(FDTD2D_ENOMEM, "out of host memory");
  int rc = fdtd2d::setup_from_ini(ini_path, s->s, objects != 0);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return FDTD2D_OK;
}

int fdtd2d_setup_ini_status(const fdtd2d_setup *s) { return s ? s->s.ini_status : 0; }

int fdtd2d_ini_lookup(const char *text, const char *section, const char *name, int kind, double def,
                      double *value, int *parse_status) {
  if (!text || !section || !name || !value) return set_error(FDTD2D_EINVAL, "null argument");
  fdtd2d::IniFile f = fdtd2d::IniFile::from_text(text);
  if (parse_status) *parse_status = f.status();
  switch (kind) {
    case 0: *value = static_cast<double>(f.get_int(section, name, static_cast<long>(def))); break;
    case 1: *value = f.get_real(section, name, def); break;
    case 2: *value = f.get_bool(section, name, def != 0.0) ? 1.0 : 0.0; break;
    default: return set_error(FDTD2D_EINVAL, "kind must be 0 (int), 1 (real) or 2 (bool)");
  }
  return FDTD2D_OK;
}

int fdtd2d_setup_workload(int workload, uint64_t seed, int shot, int steps_override,
                          fdtd2d_setup **out) {
  if (!out) return set_error(FDTD2D_EINVAL, "null out");
  auto *s = new (std::nothrow) fdtd2d_setup();
  if (!s) return set_error(FDTD2D_ENOMEM, "out of host memory");
  int rc;
  try {
    rc = fdtd2d::setup_workload(workload, seed, shot, steps_override, s->s);
  } catch (const std::bad_alloc &) {
    rc = set_error(FDTD2D_ENOMEM, "out of host memory building the workload");
  }
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return FDTD2D_OK;
}

int fdtd2d_setup_box(int nx, int ny, int npml, int nsteps, double ds_m, fdtd2d_setup **out) {
  if (!out) return set_error(FDTD2D_EINVAL, "null out");
  auto *s = new (std::nothrow) fdtd2d_setup();
  if (!s) return set_error(FDTD2D_ENOMEM, "out of host memory");
  int rc = fdtd2d::setup_box(nx, ny, npml, nsteps, ds_m, s->s);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return FDTD2D_OK;
}

const fdtd2d_processing *fdtd2d_setup_processing(const fdtd2d_setup *s) { return s ? &s->s.proc : nullptr; }

int fdtd2d_setup_set(fdtd2d_setup *s, const char *name, double value) {
  if (!s || !name) return set_error(FDTD2D_EINVAL, "null argument");
  if (std::string(name) != "SlicesSaved") return set_error(FDTD2D_EINVAL, std::string("not settable: ") + name);
  s->s.params[name] = value;
  return FDTD2D_OK;
}

const fdtd2d_problem *fdtd2d_setup_problem(const fdtd2d_setup *s) { return s ? &s->s.p : nullptr; }
fdtd2d_problem *fdtd2d_setup_problem_mut(fdtd2d_setup *s) { return s ? &s->s.p : nullptr; }

int fdtd2d_setup_get(const fdtd2d_setup *s, const char *name, double *value) {
  if (!s || !name || !value) return set_error(FDTD2D_EINVAL, "null argument");
  auto it = s->s.params.find(name);
  if (it == s->s.params.end()) return set_error(FDTD2D_EINVAL, std::string("unknown parameter ") + name);
  *value = it->second;
  return FDTD2D_OK;
}

void fdtd2d_setup_destroy(fdtd2d_setup *s) { delete s; }

}  // extern "C"
