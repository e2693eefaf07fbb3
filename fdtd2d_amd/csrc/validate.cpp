// validate.cpp - consistency checks on an fdtd2d_problem before any device
// work (the reference performs none; see F/kernel.cu:1087-1618).
#include "validate.h"

#include <string>

#include "errors.h"

namespace fdtd2d {

int validate_problem(const fdtd2d_problem *p) {
  if (!p) return set_error(FDTD2D_EINVAL, "null problem");
  if (p->npml < 0 || p->npml == 1) return set_error(FDTD2D_EINVAL, "npml must be 0 or >= 2");
  if (p->nx < 2 * p->npml + 4 || p->ny < 2 * p->npml + 4)
    return set_error(FDTD2D_EINVAL, "grid smaller than 2*npml+4 in x or y");
  if (static_cast<long long>(p->nx) * p->ny > (1LL << 31) - 1)
    return set_error(FDTD2D_EINVAL, "grid larger than 2^31 cells");
  if (p->nsteps < 0) return set_error(FDTD2D_EINVAL, "nsteps < 0");
  if (!(p->dt > 0) || !(p->dx > 0) || !(p->dy > 0)) return set_error(FDTD2D_EINVAL, "dt, dx, dy must be > 0");
  if (!p->den_ex || !p->den_hx || !p->den_ey || !p->den_hy)
    return set_error(FDTD2D_EINVAL, "missing den_* tables");
  if (p->npml > 0)
    for (int s = 0; s < 2; ++s)
      if (!p->be_x[s] || !p->ce_x[s] || !p->bh_x[s] || !p->ch_x[s] || !p->be_y[s] || !p->ce_y[s] ||
          !p->bh_y[s] || !p->ch_y[s])
        return set_error(FDTD2D_EINVAL, "missing CPML tables");
  if (p->medium_kind == FDTD2D_MEDIUM_TABLE) {
    if (!p->material || !p->ca || !p->cb || !p->ne || p->nmaterials <= 0)
      return set_error(FDTD2D_EINVAL, "TABLE medium needs material, ca, cb, ne");
    if (p->nmaterials > 256) return set_error(FDTD2D_EINVAL, "at most 256 materials");
    const size_t n = static_cast<size_t>(p->nx) * p->ny;
    for (size_t k = 0; k < n; ++k)
      if (p->material[k] < 0 || p->material[k] >= p->nmaterials)
        return set_error(FDTD2D_EINVAL, "material id out of range at cell " + std::to_string(k));
  } else if (p->medium_kind == FDTD2D_MEDIUM_DENSITY) {
    if (!p->density) return set_error(FDTD2D_EINVAL, "DENSITY medium needs density");
  } else {
    return set_error(FDTD2D_EINVAL, "unknown medium_kind");
  }
  if (p->nsrc < 0 || p->ndet < 0 || p->ncomp < 0) return set_error(FDTD2D_EINVAL, "negative counts");
  if (p->nsrc > 0 && (!p->src_i || !p->src_j || !p->src_table))
    return set_error(FDTD2D_EINVAL, "sources need src_i, src_j, src_table");
  if (p->ndet > 0 && (!p->det_i || !p->det_j)) return set_error(FDTD2D_EINVAL, "detectors need det_i, det_j");
  if (p->ncomp > 0 && (!p->comp_i || !p->comp_j))
    return set_error(FDTD2D_EINVAL, "component detectors need comp_i, comp_j");
  for (int e = 0; e < p->nsrc; ++e)
    if (p->src_i[e] < 0 || p->src_i[e] >= p->nx || p->src_j[e] < 0 || p->src_j[e] >= p->ny)
      return set_error(FDTD2D_EINVAL, "source " + std::to_string(e) + " outside the grid");
  for (int d = 0; d < p->ndet; ++d)
    if (p->det_i[d] < 0 || p->det_i[d] >= p->nx || p->det_j[d] < 0 || p->det_j[d] >= p->ny)
      return set_error(FDTD2D_EINVAL, "detector " + std::to_string(d) + " outside the grid");
  for (int c = 0; c < p->ncomp; ++c)
    if (p->comp_i[c] < 0 || p->comp_i[c] > p->nx - 2 || p->comp_j[c] < 0 || p->comp_j[c] > p->ny - 2)
      return set_error(FDTD2D_EINVAL, "component detector " + std::to_string(c) +
                                          " outside [0,nx-2]x[0,ny-2]");
  if (p->nrc < 0) return set_error(FDTD2D_EINVAL, "negative RC line count");
  if (p->nrc > 0) {
    if (!p->rc_i || !p->rc_imax) return set_error(FDTD2D_EINVAL, "RC lines need rc_i, rc_imax");
    if (p->nrc > 4096) return set_error(FDTD2D_EINVAL, "at most 4096 RC lines");
    for (int e = 0; e < p->nrc; ++e)
      if (p->rc_i[e] < 0 || p->rc_i[e] >= p->nx || p->rc_imax[e] < 0 || p->rc_imax[e] >= p->nx)
        return set_error(FDTD2D_EINVAL, "RC line " + std::to_string(e) + " outside the grid");
    if (p->rc_jcut < 0) return set_error(FDTD2D_EINVAL, "rc_jcut < 0");
  }
  if (p->line_size < 0) return set_error(FDTD2D_EINVAL, "negative processing line size");
  if (p->line_size > 0) {
    if (p->line_i0 < 0 || static_cast<long long>(p->line_i0) + p->line_size > p->nx || p->line_j < 0 ||
        p->line_j >= p->ny)
      return set_error(FDTD2D_EINVAL, "processing line outside the grid");
    if (p->line_begin < 0) return set_error(FDTD2D_EINVAL, "processing line_begin < 0");
  }
  return FDTD2D_OK;
}

}  // namespace fdtd2d
