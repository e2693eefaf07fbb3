// geometry.h - metal geometry of Settings.ini [KameraConf] written into the
// material grid (ID3, Imax x Jmax, row-major by J).
#ifndef FDTD2D_GEOMETRY_H
#define FDTD2D_GEOMETRY_H

#include <cstdint>
#include <vector>

namespace fdtd2d {

// draw_2D_figure (F/FDTD_cuda.cu:4309-4462): presets 1-7 of steel (material
// 5) walls, pipes and rings with their openings cut back to vacuum; other
// presets draw nothing. srcI is the scaled source centre I.
void draw_chamber_figure(int preset, int patL, int srcI, int nx, int ny, std::vector<int32_t> &id3);
// buildPlateMedium (F/FDTD_cuda.cu:3077-3112): steel from row wallPos down.
void build_plate_wall(int wallPos, int nx, int ny, std::vector<int32_t> &id3);

}  // namespace fdtd2d
#endif
