// validate.h - problem consistency checks shared by plan creation and the
// footprint query.
#ifndef FDTD2D_VALIDATE_H
#define FDTD2D_VALIDATE_H

#include "../../include/fdtd2d.h"

namespace fdtd2d {
int validate_problem(const fdtd2d_problem *p);
}

#endif
