// geometry.cpp - the reference's metal geometry written into the material
// grid (ID3): the 2D chamber figures of [KameraConf] Draw2D / 2Dpreset /
// PatrubokL (draw_2D_figure, F/FDTD_cuda.cu:4309-4462) and the plate wall of
// DrawWall / WallPos (buildPlateMedium, :3077-3112).
//
// Coordinates are cells (the reference does not scale them by ds), x = I
// (column), y = J (row). Every primitive marks whole Yee cells: the four Ez
// nodes (I,J), (I+1,J), (I+1,J+1), (I,J+1) (yeeCube, :2771-2790), addressed
// as the linear index J*Imax + I of the Imax x Jmax grid, so a node at
// I + 1 == Imax lands on column 0 of the next row, as in the reference.
// Nodes whose linear index falls outside the grid, which the reference would
// write past its array, are skipped.
#include <cmath>
#include <cstdint>
#include <vector>

#include "geometry.h"

namespace fdtd2d {
namespace {

struct Grid {
  int nx, ny;
  std::vector<int32_t> &id3;
  void node(long long idx, int m) {
    if (idx >= 0 && idx < static_cast<long long>(nx) * ny) id3[idx] = m;
  }
  // yeeCube(I, J, 0, m)
  void cell(int I, int J, int m) {
    const long long r0 = static_cast<long long>(J) * nx, r1 = static_cast<long long>(J + 1) * nx;
    node(r0 + I, m);
    node(r0 + I + 1, m);
    node(r1 + I + 1, m);
    node(r1 + I, m);
  }
};

// draw_line, F/FDTD_cuda.cu:4491-4523: one cell per row between y1 and y2
// (y2 excluded) at the truncated-division x, or along the row y2 from x1 to
// x2 (x2 excluded) when y1 == y2.
void line(Grid &g, int x1, int y1, int x2, int y2, int m) {
  if (y2 != y1) {
    const int d = y2 > y1 ? 1 : -1;
    for (int y = y1; y != y2; y += d) g.cell(((x2 - x1) * (y - y1)) / (y2 - y1) + x1, y, m);
  } else {
    const int d = x2 > x1 ? 1 : -1;
    for (int x = x1; x != x2; x += d) g.cell(x, y2, m);
  }
}

// draw_width_line, :4464-4489: the line and a copy one row below (segments
// going up-right or down-left: dx, dy of the same sign) or above; a segment
// with dx == 0 or dy == 0 matches two of the four sign tests and is drawn
// again for each.
void wide_line(Grid &g, int x1, int y1, int x2, int y2, int m) {
  const int dx = x2 - x1, dy = y2 - y1;
  if (dx >= 0 && dy >= 0) {
    line(g, x1, y1, x2, y2, m);
    line(g, x1, y1 - 1, x2, y2 - 1, m);
  }
  if (dx <= 0 && dy <= 0) {
    line(g, x1, y1, x2, y2, m);
    line(g, x1, y1 + 1, x2, y2 + 1, m);
  }
  if (dx <= 0 && dy >= 0) {
    line(g, x1, y1, x2, y2, m);
    line(g, x1, y1 + 1, x2, y2 + 1, m);
  }
  if (dx >= 0 && dy <= 0) {
    line(g, x1, y1, x2, y2, m);
    line(g, x1, y1 - 1, x2, y2 - 1, m);
  }
}

// draw_Circl, :4525-4546: cells (x, y), x < Imax-1, y < Jmax-1, with
// (x-a)^2 + (y-b)^2 <= R^2 take m, then those within R - width take 0.
void ring(Grid &g, int a, int b, int R, int width, int m) {
  for (int pass = 0; pass < 2; ++pass) {
    const double rr = pass == 0 ? double(R) * R : double(R - width) * (R - width);
    for (int x = 0; x < g.nx - 1; ++x)
      for (int y = 0; y < g.ny - 1; ++y) {
        const double d = double(x - a) * (x - a) + double(y - b) * (y - b);
        if (d <= rr) g.cell(x, y, pass == 0 ? m : 0);
      }
  }
}

// Draw_Square, :4549-4564: cells strictly inside (x1, x2) x (y1, y2).
void square(Grid &g, int x1, int y1, int x2, int y2, int m) {
  for (int x = 0; x < g.nx - 1; ++x)
    for (int y = 0; y < g.ny - 1; ++y)
      if (x > x1 && x < x2 && y > y1 && y < y2) g.cell(x, y, m);
}

}  // namespace

void draw_chamber_figure(int preset, int patL, int srcI, int nx, int ny, std::vector<int32_t> &id3) {
  Grid g{nx, ny, id3};
  const int Imax = nx, Jmax = ny;
  const int RB = 270, RS = 60, PX = 100, PY = patL;  // radii, pipe width, pipe length
  const int m1 = 5, m2 = 5;                          // steel
  const int kISource = 300, kAaSzI = 40;             // I_SOURCE, AA_SZ_I (F/Settings.cuh:30, 44)
  const int top = 180 + PY;
  auto pipe = [&] {
    wide_line(g, 600, top, 600, 180, m2);
    wide_line(g, 600, 180, 1000, 180, m2);
    wide_line(g, 1000, 180, 1000, top, m2);
    wide_line(g, 600, top, 1000, top, 0);
  };
  switch (preset) {
    case 1:  // chamber with an antenna hole
      wide_line(g, 20, 180, Imax - 20, 180, m1);
      wide_line(g, Imax - 20, 180, Imax - 20, Jmax - 20, m1);
      wide_line(g, Imax - 20, Jmax - 20, 20, Jmax - 20, m1);
      wide_line(g, 20, Jmax - 20, 20, 180, m1);
      wide_line(g, kISource - 25, 180, kISource + 25, 180, 0);
      break;
    case 2:  // chamber behind a pipe, antenna opening cut through the front
      wide_line(g, 20, top, Imax - 20, top, m1);
      wide_line(g, Imax - 20, top, Imax - 20, Jmax - 20, m1);
      wide_line(g, Imax - 20, Jmax - 20, 20, Jmax - 20, m1);
      wide_line(g, 20, Jmax - 20, 20, top, m1);
      pipe();
      square(g, srcI - kAaSzI / 2, 11, srcI + kAaSzI / 2, 220, 0);
      break;
    case 3:  // no back wall
      wide_line(g, 20, top, Imax - 20, top, m1);
      wide_line(g, Imax - 20, top, Imax - 20, Jmax - 20, m1);
      wide_line(g, 20, Jmax - 20, 20, top, m1);
      pipe();
      wide_line(g, kISource - 25, 180, kISource + 25, 180, 0);
      break;
    case 4:  // front wall only
      wide_line(g, 20, top, Imax - 20, top, m1);
      pipe();
      wide_line(g, kISource - 25, 180, kISource + 25, 180, 0);
      break;
    case 5:  // the pipe alone
      pipe();
      wide_line(g, kISource - 25, 180, kISource + 25, 180, 0);
      break;
    case 6:  // front and back walls
      wide_line(g, 20, top, Imax - 20, top, m1);
      wide_line(g, Imax - 20, Jmax - 20, 20, Jmax - 20, m1);
      pipe();
      wide_line(g, kISource - 25, 180, kISource + 25, 180, 0);
      break;
    case 7: {  // tokamak equatorial section: two rings, a port and its pipe
      const int a = Imax / 2, b = Jmax / 2 + PY;
      ring(g, a, b, RB, 3, m1);
      ring(g, a, b, RS, 4, m1);
      square(g, a - PX / 2, 11, a + PX / 2, Jmax / 2 + PY + 15 - RB, 0);
      wide_line(g, a - PX / 2, Jmax / 2 + PY + 4 - RB, a - PX / 2, Jmax / 2 - PY - RB, m2);
      wide_line(g, a - PX / 2, Jmax / 2 - PY - RB, a + PX / 2, Jmax / 2 - PY - RB, m2);
      wide_line(g, a + PX / 2, Jmax / 2 - PY - RB, a + PX / 2, Jmax / 2 + PY + 4 - RB, m2);
      square(g, srcI - kAaSzI / 2, 11, srcI + kAaSzI / 2, Jmax / 2 + PY + 4 - RB, 0);
      break;
    }
    default:
      break;
  }
}

void build_plate_wall(int wallPos, int nx, int ny, std::vector<int32_t> &id3) {
  Grid g{nx, ny, id3};
  for (int i = 1; i < nx - 1; ++i)
    for (int j = wallPos; j < ny - 1; ++j) g.cell(i, j, 5);
}

}  // namespace fdtd2d
