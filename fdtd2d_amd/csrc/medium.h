// medium.h - the reference's plasma medium builders: Settings.ini [medium]
// and [mediumFluc] keys -> material ids in ID3 (F/FDTD_cuda.cu:2602-2659,
// 2793-2972, 3114-3378).
#ifndef FDTD2D_MEDIUM_H
#define FDTD2D_MEDIUM_H

#include <cstdint>
#include <vector>

namespace fdtd2d {

class IniFile;

struct MediumParams {
  int build = 0;  // 0 none, 1 sphere/ellipse, 2 torus
  int cenI = 275, cenJ = 305, cenK = 1;
  float A = 1, B = 1, C = 1;
  int P = 1, Q = 1, N = 1;
  int R = 50;
  float alpha = 2.0f, beta = 1.0f, ne0 = 1e19f, neA = 0;
  int torusR = 175;
  int fluc_build = 0;
  int fcenI = 275, fcenJ = 305, fcenK = 1;
  float fA = 1, fB = 1, fC = 1;
  int fP = 1, fQ = 1, fN = 1;
  float angI = 0, angJ = 0, angK = 0;
  int sign = 1;
  int fR = 10;
  float falpha = 2.0f, fbeta = 1.0f, fne0 = 1e19f;
};

void read_medium_settings(const IniFile &r, MediumParams &m);
// recuclParametersToMetric's scaling of the medium keys (F/FDTD_cuda.cu:327-363).
void scale_medium_settings(MediumParams &m, double ds);
// Writes material ids (ID3) for an Imax x Jmax grid; ne is the pne table.
int build_reference_medium(const MediumParams &m, int nx, int ny, const std::vector<float> &ne,
                           std::vector<int32_t> &id3);

}  // namespace fdtd2d
#endif
