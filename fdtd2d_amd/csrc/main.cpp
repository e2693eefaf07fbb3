// main.cpp - command-line host driver, the role of FDTD_cuda.cu's main() /
// initlimits() (F/FDTD_cuda.cu:20-23, 384-441): read Settings.ini from the
// working directory (or the path given), build the problem, run it on the
// GPU named by [hardware] GPU_device, write the reference's output files
// (detectors, Ez snapshots, RC tracker, processing line, and with
// [regimes] SourceToDetect = 1 the free-space pass and the difference).
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

struct Traces {
  std::vector<float> det, comp;
};

// Runs n steps of the plan, saving Ez after every snapshot step when w is
// given: the launches up to the next snapshot are queued before the previous
// snapshot is formatted, so the text formatting overlaps the device work.
static int run_steps(fdtd2d_plan *plan, const fdtd2d_setup *s, int n, fdtd2d_snapshots *w) {
  int rc;
  std::vector<int32_t> snaps;
  if (w) {
    size_t c = 0;
    if ((rc = fdtd2d_snapshot_steps(s, n, nullptr, 0, &c))) return rc;
    snaps.resize(c);
    if (c && (rc = fdtd2d_snapshot_steps(s, n, snaps.data(), c, &c))) return rc;
  }
  const fdtd2d_problem *p = fdtd2d_setup_problem(s);
  const size_t cells = static_cast<size_t>(p->nx) * p->ny;
  std::vector<float> buf[2];
  int done = 0, pending = -1;
  for (size_t k = 0; k < snaps.size(); ++k) {
    if ((rc = fdtd2d_run(plan, snaps[k] - done))) return rc;
    done = snaps[k];
    if (pending >= 0 && (rc = fdtd2d_snapshots_add(w, buf[pending].data(), cells))) return rc;
    const int b = static_cast<int>(k & 1);
    buf[b].resize(cells);
    if ((rc = fdtd2d_read_field(plan, FDTD2D_EZ, buf[b].data(), cells))) return rc;
    pending = b;
  }
  if ((rc = fdtd2d_run(plan, n - done))) return rc;
  if (pending >= 0 && (rc = fdtd2d_snapshots_add(w, buf[pending].data(), cells))) return rc;
  return fdtd2d_sync(plan);
}

static int read_traces(fdtd2d_plan *plan, const fdtd2d_problem *p, Traces &t) {
  t.det.assign(static_cast<size_t>(p->nsteps + 1) * p->ndet, 0.f);
  t.comp.assign(static_cast<size_t>(3) * (p->nsteps + 1) * p->ncomp, 0.f);
  return fdtd2d_read_detectors(plan, t.det.data(), t.det.size(), t.comp.data(), t.comp.size());
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
  if ((rc = fdtd2d_write_ini_log(s, out.c_str()))) return die(rc, "write_ini_log");
  const fdtd2d_problem *p = fdtd2d_setup_problem(s);
  double gpu = 0, s2d = 0;
  fdtd2d_setup_get(s, "GPU_device", &gpu);
  fdtd2d_setup_get(s, "SourceToDetect", &s2d);
  if (device < 0) device = static_cast<int>(gpu);
  const int n = steps >= 0 && steps < p->nsteps ? steps : p->nsteps;
  std::printf("\nTIme step = %e\n Number of steps = %d\n Total Simulation time = %e Seconds\n", p->dt,
              p->nsteps, p->nsteps * p->dt);

  // Source-to-detector regime: a first pass without objects (no chamber
  // figure, no plasma), whose traces the main pass's are compared with.
  Traces free_space;
  if (s2d == 1) {
    std::printf("\n\nStart Source To Detect Influence Calculation\n");
    fdtd2d_setup *s0 = nullptr;
    if ((rc = fdtd2d_setup_from_ini_objects(ini.c_str(), 0, &s0))) return die(rc, "setup (free space)");
    fdtd2d_problem *p0 = fdtd2d_setup_problem_mut(s0);
    p0->nrc = 0;  // that pass runs neither the RC tracker nor the processing line
    p0->line_size = 0;
    fdtd2d_plan *plan0 = nullptr;
    if ((rc = fdtd2d_plan_create(p0, device, &plan0))) return die(rc, "plan_create (free space)");
    if (tblock && (rc = fdtd2d_plan_set_tblock(plan0, tblock))) return die(rc, "set_tblock");
    if ((rc = run_steps(plan0, s0, n, nullptr))) return die(rc, "run (free space)");
    if ((rc = read_traces(plan0, p0, free_space))) return die(rc, "read_detectors (free space)");
    if ((rc = fdtd2d_write_s2d_outputs(s0, out.c_str(), free_space.det.data(), free_space.comp.data())))
      return die(rc, "write_s2d_outputs");
    fdtd2d_plan_destroy(plan0);
    fdtd2d_setup_destroy(s0);
    std::printf("Done Source To Detect Influence Calculation\n\n");
  }

  auto t0 = std::chrono::steady_clock::now();
  fdtd2d_plan *plan = nullptr;
  if ((rc = fdtd2d_plan_create(p, device, &plan))) return die(rc, "plan_create");
  if (tblock && (rc = fdtd2d_plan_set_tblock(plan, tblock))) return die(rc, "set_tblock");
  fdtd2d_snapshots *w = nullptr;
  if ((rc = fdtd2d_snapshots_open(s, out.c_str(), &w))) return die(rc, "snapshots_open");
  if ((rc = run_steps(plan, s, n, w))) return die(rc, "run");
  int saved = 0;
  if ((rc = fdtd2d_snapshots_close(w, &saved))) return die(rc, "snapshots_close");
  fdtd2d_setup_set(s, "SlicesSaved", saved);
  Traces tr;
  if ((rc = read_traces(plan, p, tr))) return die(rc, "read_detectors");
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("Done time-stepping...\n");
  if ((rc = fdtd2d_write_outputs(s, out.c_str(), tr.det.data(), tr.comp.data(), secs)))
    return die(rc, "write_outputs");
  if (s2d == 1 && (rc = fdtd2d_write_s2d_difference(s, out.c_str(), tr.comp.data(), free_space.comp.data())))
    return die(rc, "write_s2d_difference");
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
    if (times == 0) {
      std::printf("Processing skipped: no step of its window (%d, %d] was run\n", a->begin_after, a->finish_at);
    } else {
      std::printf("Processing...\n");
      if ((rc = fdtd2d_write_processing_outputs(a, out.c_str(), line.data(), times)))
        return die(rc, "processing");
      std::printf("Processing done.\n");
    }
  }
  std::printf("%d steps of %dx%d in %.3f s, %d Ez snapshots\n", n, p->nx, p->ny, secs, saved);
  fdtd2d_plan_destroy(plan);
  fdtd2d_setup_destroy(s);  // p points into s
  return 0;
}
