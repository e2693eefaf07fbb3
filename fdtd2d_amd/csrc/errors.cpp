// errors.cpp - status strings and the per-thread last error of the C ABI.
#include "errors.h"

#include "../../include/fdtd2d.h"

namespace fdtd2d {
namespace {
thread_local std::string g_last;
}
int set_error(int code, const std::string &msg) {
  g_last = msg;
  return code;
}
const char *last_error() { return g_last.c_str(); }
}  // namespace fdtd2d

extern "C" {

const char *fdtd2d_strerror(int status) {
  switch (status) {
    case FDTD2D_OK: return "ok";
    case FDTD2D_EINVAL: return "invalid argument";
    case FDTD2D_ENOMEM: return "out of memory";
    case FDTD2D_EDEVICE: return "device error";
    case FDTD2D_EIO: return "i/o error";
    case FDTD2D_ESTATE: return "invalid state";
    case FDTD2D_EUNSUPPORTED: return "unsupported";
  }
  return "unknown status";
}

const char *fdtd2d_last_error(void) { return fdtd2d::last_error(); }

int fdtd2d_abi_version(void) { return FDTD2D_ABI_VERSION; }
}
