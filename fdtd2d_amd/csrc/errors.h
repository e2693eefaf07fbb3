// errors.h - thread-local last-error text behind the int status codes.
#ifndef FDTD2D_ERRORS_H
#define FDTD2D_ERRORS_H

#include <string>

namespace fdtd2d {
// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string &msg);
const char *last_error();
}  // namespace fdtd2d

#endif
