// main.cpp - command-line host driver, the role of FDTD_cuda.cu's main() /
// initlimits() (F/FDTD_cuda.cu:20-23, 384-441): read Settings.ini from the
// working directory (or the path given), build the problem, run it on the
// GPU named by [hardware] GPU_device, write the reference's output files
// (detectors, RC tracker, processing line).
//
//   fdtd2d [Settings.ini] [--out DIR] [--steps N] [--device D] [--tblock T]
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fdtd2d.h"

static int die(int rc, const char *what) {
  std::fprintf(stderr, "fdtd2d: %s: %s (%s)\n", what, fdtd2d_strerror(rc), fdtd2d_last_error());
  return 1;
}

int main(int argc, char **argv) {
  std::string ini = "Settings.ini", out = ".";
  int steps = -1, device = -1, tblock = 0;
  for (int k = 1; k < argc; ++k) {
    std::string a = argv[k];
    if (a == "--out" && k + 1 < argc) out = argv[++k];
    else if (a == "--steps" && k + 1 < argc) steps = std::atoi(argv[++k]);
    else if (a == "--device" && k + 1 < argc) device = std::atoi(argv[++k]);
    else if (a == "--tblock" && k + 1 < argc) tblock = std::atoi(argv[++k]);
    else if (a == "-h" || a == "--help") {
      std::printf("usage: fdtd2d [Settings.ini] [--out DIR] [--steps N] [--device D] [--tblock T]\n");
      return 0;
    } else ini = a;
  }
  fdtd2d_setup *s = nullptr;
  int rc = fdtd2d_setup_from_ini(ini.c_str(), &s);
  if (rc) return die(rc, "setup");
  if (fdtd2d_setup_ini_status(s) < 0) std::printf("Can't load '%s'\n", ini.c_str());
  const fdtd2d_problem *p = fdtd2d_setup_problem(s);
  double gpu = 0;
  fdtd2d_setup_get(s, "GPU_device", &gpu);
  if (device < 0) device = static_cast<int>(gpu);
  const int n = steps >= 0 && steps < p->nsteps ? steps : p->nsteps;
  std::printf("\nTIme step = %e\n Number of steps = %d\n Total Simulation time = %e Seconds\n", p->dt,
              p->nsteps, p->nsteps * p->dt);
  auto t0 = std::chrono::steady_clock::now();
  fdtd2d_plan *plan = nullptr;
  if ((rc = fdtd2d_plan_create(p, device, &plan))) return die(rc, "plan_create");
  if (tblock && (rc = fdtd2d_plan_set_tblock(plan, tblock))) return die(rc, "set_tblock");
  if ((rc = fdtd2d_run(plan, n))) return die(rc, "run");
  if ((rc = fdtd2d_sync(plan))) return die(rc, "sync");
  std::vector<float> det(static_cast<size_t>(p->nsteps + 1) * p->ndet);
  std::vector<float> comp(static_cast<size_t>(3) * (p->nsteps + 1) * p->ncomp);
  if ((rc = fdtd2d_read_detectors(plan, det.data(), det.size(), comp.data(), comp.size())))
    return die(rc, "read_detectors");
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Done time-stepping...\n");
  if ((rc = fdtd2d_write_outputs(s, out.c_str(), det.data(), comp.data(), secs))) return die(rc, "write_outputs");
  {
    std::vector<int32_t> spd(static_cast<size_t>(p->nsteps + 1) * p->nrc);
    if (p->nrc > 0 && (rc = fdtd2d_read_rc(plan, spd.data(), spd.size()))) return die(rc, "read_rc");
    if ((rc = fdtd2d_write_rc_outputs(s, out.c_str(), spd.data()))) return die(rc, "write_rc_outputs");
  }
  // processing_core on the captured line (F/kernel.cu:2313-2315), over the
  // steps of its window this run reached
  const fdtd2d_processing *a = fdtd2d_setup_processing(s);
  if (a->perform) {
    const int times = p->line_size > 0 ? std::max(0, std::min(p->line_end, n) - p->line_begin) : 0;
    const size_t rows = p->line_size > 0 ? static_cast<size_t>(std::max(0, std::min(p->line_end, p->nsteps) -
                                                                        p->line_begin))
                                          : 0;
    std::vector<float> line(rows * static_cast<size_t>(p->line_size > 0 ? p->line_size : 0));
    if ((rc = fdtd2d_read_line(plan, line.data(), line.size()))) return die(rc, "read_line");
    std::printf("Processing...\n");
    if ((rc = fdtd2d_write_processing_outputs(a, out.c_str(), line.data(), times)))
      return die(rc, "processing");
    std::printf("Processing done.\n");
  }
  std::printf("%d steps of %dx%d in %.3f s\n", n, p->nx, p->ny, secs);
  fdtd2d_plan_destroy(plan);
  fdtd2d_setup_destroy(s);  // p points into s
  return 0;
}
