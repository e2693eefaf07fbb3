// medium.cpp - plasma medium builders of the reference (see medium.h).
#include "medium.h"

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

int build_reference_medium(const MediumParams &m, int nx, int ny, const std::vector<float> &ne,
                           std::vector<int32_t> &id3) {
  (void)nx;
  (void)ny;
  (void)ne;
  (void)id3;
  if (m.build != 0 || m.fluc_build != 0)
    return set_error(FDTD2D_EUNSUPPORTED, "Settings.ini: plasma medium builders not built yet");
  return FDTD2D_OK;
}

}  // namespace fdtd2d
