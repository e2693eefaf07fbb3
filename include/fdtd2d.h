/*
 * fdtd2d.h - C ABI of the MI355X (gfx950) 2D TMz FDTD time-stepper.
 *
 * Plain C: fixed-width integers, plain pointers and sizes, no HIP or torch
 * types. Every call returns an int status (FDTD2D_OK == 0); the text of the
 * last failure on the calling thread is available from fdtd2d_last_error().
 *
 * Reference interface each entry point replaces (Virtuoso2598/FDTD_2D,
 * FDTD_2D_Full/ abbreviated F/):
 *   the whole device side of one run is DevFieldCalculation(void)
 *   (F/Kernel.cuh:102, F/kernel.cu:1928), which reads host globals
 *   (F/Ext_kernel.cuh:8-258) and writes ~22 text files. Here the globals
 *   become an fdtd2d_problem, device state lives in an opaque plan, and file
 *   writing stays in the host driver (fdtd2d_write_outputs).
 *
 * Field layout (the reference's own index rules, F/kernel.cu:127-537):
 *   Ez, Jz : ny rows x nx columns, element (j,i) at j*nx+i
 *   Hx     : (ny-1) x nx        (pitch nx)
 *   Hy     : ny x (nx-1)        (pitch nx-1)
 * nx == Imax, ny == Jmax. The device pads every field to a common pitch; the
 * read/write calls below convert to and from these reference layouts.
 */
#ifndef FDTD2D_H
#define FDTD2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDTD2D_ABI_VERSION 4

/* status codes */
enum {
  FDTD2D_OK = 0,
  FDTD2D_EINVAL = 1,      /* bad argument or inconsistent problem */
  FDTD2D_ENOMEM = 2,      /* host or device allocation failed / over budget */
  FDTD2D_EDEVICE = 3,     /* HIP runtime error (launch, copy, fault at sync) */
  FDTD2D_EIO = 4,         /* file could not be read or written */
  FDTD2D_ESTATE = 5,      /* call not valid in the object's current state */
  FDTD2D_EUNSUPPORTED = 6 /* valid request this build does not implement */
};

/* medium_kind */
enum {
  FDTD2D_MEDIUM_TABLE = 0,  /* int32 material id per cell + CA/CB/Ne tables (reference ID3) */
  FDTD2D_MEDIUM_DENSITY = 1 /* float electron density per cell, uniform CA/CB */
};

/* fields for fdtd2d_read_field / fdtd2d_write_field */
enum { FDTD2D_EZ = 0, FDTD2D_HX = 1, FDTD2D_HY = 2, FDTD2D_JZ = 3 };

/* CPML strip arrays for fdtd2d_read_psi / fdtd2d_write_psi */
enum { FDTD2D_PSI_HXY = 0, FDTD2D_PSI_HYX = 1, FDTD2D_PSI_EZX = 2, FDTD2D_PSI_EZY = 3 };

/*
 * One simulation. Caller-owned; fdtd2d_plan_create copies everything it needs.
 *
 * Step n (1..nsteps) performs, in this order (F/kernel.cu:2073-2157):
 *   Hx, then its y-CPML psi; Hy, then its x-CPML psi; Jz from E^n;
 *   Ez (+x-CPML psi, then +y-CPML psi); soft sources Ez -= src_table[n][e];
 *   detector and component sampling into row n.
 */
typedef struct fdtd2d_problem {
  int32_t nx, ny;   /* Imax, Jmax (cells); both >= 2*npml + 4 */
  int32_t npml;     /* CPML thickness P (reference PML_SIZE = 11); 0 = bare PEC box */
  int32_t nsteps;   /* nMax: rows 0..nsteps of src_table and of the traces */
  double dt, dx, dy;

  /* H update: Hx += db*((Ez[j]-Ez[j+1])*den_hy[j]); reference DA == 1, DB == dt/mu0
     for every material (F/FDTD_cuda.cu:521-533, 1196-1199). */
  float db;

  /* E update and plasma current. */
  int32_t medium_kind;
  const int32_t *material; /* TABLE: [ny*nx] id, reference ID3 */
  int32_t nmaterials;
  const float *ca, *cb, *ne; /* TABLE: [nmaterials] CA, CB and electron density (pne) */
  int32_t pec_material;      /* TABLE: id whose Ez is forced to 0 (reference: 1); -1 = none */
  const float *density;      /* DENSITY: [ny*nx] electron density, m^-3 */
  float ca0, cb0;            /* DENSITY: CA, CB of every cell */
  float cj;                  /* e^2/m_e as float, reference D_CJ = 2.81794e-8 */
  float jz_decay;            /* Jz <- jz_decay*Jz + ((cj*Ez)*Ne)*jz_dt where Ne > 0 */
  double jz_dt;              /* collisionless reference: jz_decay = 1, jz_dt = dt */

  /* CPML. Index 0 = low side, 1 = high side; arrays are indexed by the distance
     from the outer wall exactly as the reference's be_x_1/be_x_2 etc. */
  const float *be_x[2], *ce_x[2]; /* [npml]   (index 0 unused by the updates) */
  const float *bh_x[2], *ch_x[2]; /* [npml-1] */
  const float *be_y[2], *ce_y[2]; /* [npml]   */
  const float *bh_y[2], *ch_y[2]; /* [npml-1] */
  const float *den_ex, *den_hx;   /* [nx-1] */
  const float *den_ey, *den_hy;   /* [ny-1] */
  float dx_f, dy_f;               /* psi divisors, as the float kernel argument */

  /* Soft sources (reference AESA source, F/kernel.cu:621-659). */
  int32_t nsrc;
  const int32_t *src_i, *src_j;  /* cell of element e */
  const float *src_table;        /* [(nsteps+1)*nsrc] value subtracted at step n */

  /* Ez detectors (F/kernel.cu:762-805): row n of the trace <- Ez after step n. */
  int32_t ndet;
  const int32_t *det_i, *det_j;

  /* Component detectors (F/kernel.cu:719-759): Ez, Hx, Hy at a cell. */
  int32_t ncomp;
  const int32_t *comp_i, *comp_j;

  /* RC cutoff tracker (Dev_RCcalc, F/kernel.cu:700-713, 968-1044, run after
     the detectors of every step, :2170), nrc lines, 0 = off. Line e keeps an
     edge level, starting at rc_edge0: each step it rises to rc_level times the
     signed maximum of Ez over row 0 and rows [rc_jcut, ny-3] of column
     rc_imax[e] when that is larger; the step's value is then the largest row
     j in [1, ny-3] of column rc_i[e] whose |Ez| is not below the level (0 when
     none). Row n of the trace <- step n; row 0 = 0. */
  int32_t nrc;
  const int32_t *rc_i, *rc_imax;
  int32_t rc_jcut;
  float rc_level, rc_edge0;

  /* Processing line (Dev_SaveLine, F/kernel.cu:666-688, copied to the host
     each step n with line_begin < n <= line_end, :2138-2145): Ez of row
     line_j, columns [line_i0, line_i0 + line_size), read after the sources of
     step n. Capture row t = n - line_begin - 1 for the
     min(line_end, nsteps) - line_begin captured steps; line_size 0 = off.
     (The reference kernel stores only when line_j == 1, its thread index
     never reaching another row; this is the line its arguments name, and
     equals the reference's capture when line_j == 1 and line_i0 == 1.) */
  int32_t line_size, line_i0, line_j, line_begin, line_end;
} fdtd2d_problem;

/* [Processing] of Settings.ini (struct Analysis, F/processing.cuh:8-51, read
 * at F/FDTD_cuda.cu:144-172, scaled at :364-377): time-of-flight search over
 * the captured line (processing_core, F/processing.cu:6-144). Positions and
 * size are in cells (ints divided by ds, truncated), begin_after / finish_at
 * in steps; times = captured steps; dt = the step (s) of the TOF columns. */
typedef struct fdtd2d_processing {
  int32_t perform;
  int32_t sec_der, sdf_type, sdf_win, sdf_runs;  /* second-derivative method and its post-filter */
  int32_t cf, cf_tau;                             /* constant fraction: delay (steps), attenuation */
  float cf_k;
  int32_t arc, arc_tau;
  float arc_k;
  float tau_threshold;                            /* fraction of the maximum a sample must exceed */
  int32_t filter_type, filter_win, filter_runs;   /* 1 moving average, 2 median, other: none */
  int32_t save_original, save_filtered, dim_sieve, time_sieve;
  int32_t sim_detect, save_detector, load_weights, detector_size;
  int32_t begin_after, finish_at, pos_i, pos_j, pos_k, size_i;
  int32_t times;
  double dt;
} fdtd2d_processing;

typedef struct fdtd2d_plan fdtd2d_plan;
typedef struct fdtd2d_setup fdtd2d_setup;

/* ------------------------------------------------------------------ errors */
const char *fdtd2d_strerror(int status);
const char *fdtd2d_last_error(void);
int fdtd2d_abi_version(void);
/* Number of visible HIP devices (0 when none; no device is initialised beyond
   the runtime's own discovery). */
int fdtd2d_device_count(int *count);

/* ------------------------------------------------------------------ setup
 * Host-side builders that own the arrays an fdtd2d_problem points to. */

/* Settings.ini -> problem, with the reference's semantics:
 * readINIparam + recuclParametersToMetric + setUp + buildObject +
 * buildSourceProfile + initializeCPML (F/FDTD_cuda.cu:27-176, 305-382,
 * 1122-1579, 2602-2659, 3380-3655). A missing or unparsable file falls back
 * to the compile-time defaults, as the reference does (F/FDTD_cuda.cu:30-33);
 * fdtd2d_setup_ini_status reports what happened (0 ok, -1 open failed, >0 first bad line). */
int fdtd2d_setup_from_ini(const char *ini_path, fdtd2d_setup **out);
/* As fdtd2d_setup_from_ini with the reference's ObjectsEnable flag
 * (F/FDTD_cuda.cu:389-419, 2602-2659): objects = 0 builds neither the chamber
 * figure nor the plasma medium, the first (free-space) pass of the
 * source-to-detector regime ([regimes] SourceToDetect = 1); objects = 1 is
 * fdtd2d_setup_from_ini. */
int fdtd2d_setup_from_ini_objects(const char *ini_path, int objects, fdtd2d_setup **out);
int fdtd2d_setup_ini_status(const fdtd2d_setup *s);

/* One Settings.ini value from INI text in memory, read as the reference's
 * INIReader does (F/INIlib/INIReader.cpp:30-60): kind 0 GetInteger
 * (strtol base 0), 1 GetReal (strtod), 2 GetBoolean; `def` when absent or
 * unparsable. *parse_status (may be NULL) gets 0 or the first bad line. */
int fdtd2d_ini_lookup(const char *text, const char *section, const char *name, int kind, double def,
                      double *value, int *parse_status);

/* Synthetic workloads, seeded (std::mt19937_64):
 *   1  400x400 plain method (one point source, linear plasma ramp, 4 detectors)
 *   2  8192x8192 single grid (16-element steered array, parabolic plasma x (1+0.02 g))
 *   3  one shot of the 256-shot uneven batch mix (shot = 0..255)
 *   4  one frequency of a served sweep request (shot = request*64 + frequency index)
 *   5  one shot of the mixed-size serving stream (shot = request index)
 * steps_override > 0 replaces the workload's step count; for workloads 3-5,
 * steps_override < 0 only makes the random draws (fdtd2d_setup_get of
 * "Imax", "Jmax", "nMax", "nfreq", "big", ...) and builds no problem. */
int fdtd2d_setup_workload(int workload, uint64_t seed, int shot, int steps_override,
                          fdtd2d_setup **out);

/* Tiny grids for tests: PEC box (npml = 0) or CPML box, uniform vacuum,
 * no sources, no detectors. */
int fdtd2d_setup_box(int nx, int ny, int npml, int nsteps, double ds_m, fdtd2d_setup **out);

const fdtd2d_problem *fdtd2d_setup_problem(const fdtd2d_setup *s);
fdtd2d_problem *fdtd2d_setup_problem_mut(fdtd2d_setup *s);
/* Scalar value of a named setup parameter (e.g. "dt", "Imax", "nMax",
 * "fCoef", "qantI"); returns FDTD2D_EINVAL for unknown names. */
int fdtd2d_setup_get(const fdtd2d_setup *s, const char *name, double *value);
/* Sets a value the writers report: "SlicesSaved" (Ez snapshots written, the
 * "Slises saved" line of log.txt). */
int fdtd2d_setup_set(fdtd2d_setup *s, const char *name, double value);
void fdtd2d_setup_destroy(fdtd2d_setup *s);

/* Reference output files (Detect.txt, detAll.txt, detEz/Hx/Hy.txt,
 * SourceP.txt, log.txt ...; F/kernel.cu:2409-2493, F/FDTD_cuda.cu:3647-3650)
 * written into dir from traces laid out as fdtd2d_read_detectors returns them. */
/* RC tracker products of the reference's post-processing, from traces laid out
 * as fdtd2d_read_rc returns them: RC_Line.txt and TAU_Line.txt
 * (CalculateRCfromPwrLines / CalculateTaufromPwrLines after the moving-average
 * filter RClineFilterPostTime, F/kernel.cu:3180-3270, 2628-2644) and
 * RcDet_data.txt (line positions, F/FDTD_cuda.cu:3492-3536). */
int fdtd2d_write_rc_outputs(const fdtd2d_setup *s, const char *dir, const int32_t *spd);
/* The setup's [Processing] parameters (never NULL for a valid setup; perform
 * = 0 when the INI has none). */
const fdtd2d_processing *fdtd2d_setup_processing(const fdtd2d_setup *s);
/* Time-of-flight search of processing_core (F/processing.cu:6-144) over a
 * captured line laid out as fdtd2d_read_line returns it ([times][size_i]).
 * pass 0: the line itself, n = size_i coordinates; pass 1: the simulated
 * detector (a weighted sum over detector_size neighbours, :64-100), n =
 * size_i - detector_size; weights (detector_size floats) may be NULL for the
 * uniform 1/detector_size. sd, cf, arc (n floats each, may be NULL) get
 * begin_after + the step index each method finds (the last index when none);
 * original and filtered ([n][times], may be NULL) the per-coordinate |Ez|
 * (or detector) series before and after the prefilter. Inputs on which the
 * reference reads outside its arrays (series shorter than a filter window or
 * a delay + 1, zero sieves or windows, detector_size >= size_i) are
 * FDTD2D_EINVAL here. */
int fdtd2d_processing_tof(const fdtd2d_processing *a, const float *line, int times, const float *weights,
                          int pass, float *sd, float *cf, float *arc, float *original, float *filtered);
/* processing_core's files, written into dir from a captured line: SD.txt,
 * CF.txt, ARC.txt, original_field.txt, filtered_field.txt and, with
 * sim_detect, det_SD/det_CF/det_ARC.txt and detector_original_field.txt /
 * detector_filtered_field.txt (F/processing.cu:44-140, 262-283); with
 * load_weights the detector weights come from dir/DetWeights.txt as the
 * reference reads its working directory's. */
int fdtd2d_write_processing_outputs(const fdtd2d_processing *a, const char *dir, const float *line, int times);
int fdtd2d_write_outputs(const fdtd2d_setup *s, const char *dir, const float *det,
                         const float *comp, double seconds);
/* iniLOG.txt (appended): the configuration echo the reference writes after
 * reading Settings.ini (logINIread, F/FDTD_cuda.cu:178-303), with the keys'
 * values as read (before the conversion to cells) and the time of the read;
 * nothing when the setup did not come from an INI file that could be opened.
 * (fdtd2d_write_outputs also writes deb.txt, the material id map
 * F/kernel.cu:2737-2751, for table-medium problems, and creates the empty
 * files the reference opens: test.txt, SDetect.txt, LineDetect.txt,
 * LineDetect2.txt, detEx/Ey/Hz.txt, DifDetEx.txt, CompDetect.txt,
 * materials.txt.) */
int fdtd2d_write_ini_log(const fdtd2d_setup *s, const char *dir);
/* Source-to-detector regime ([regimes] SourceToDetect = 1, F/FDTD_cuda.cu:
 * 389-419): the free-space pass's files (Detect_S2D.txt, detAll_S2D.txt,
 * detEz/Hx/Hy_S2D.txt and the empty files the reference opens,
 * DevSourceToDetectInfluenceCalc F/kernel.cu:1620-1926) from that pass's
 * traces, and DifDetEx.txt (F/kernel.cu:2497-2536) from the component traces
 * of the main pass minus those of the free-space pass. */
int fdtd2d_write_s2d_outputs(const fdtd2d_setup *s, const char *dir, const float *det, const float *comp);
int fdtd2d_write_s2d_difference(const fdtd2d_setup *s, const char *dir, const float *comp,
                                const float *comp_free);

/* Ez snapshots of [2Dsave] (F/kernel.cu:2173-2302, 3041-3051). The steps
 * n <= nsteps_run after which the reference saves Ez: a counter starting at 1
 * counts the steps inside [From, To] and saves when it equals TimeSieve
 * (steps may be NULL to query *count). */
int fdtd2d_snapshot_steps(const fdtd2d_setup *s, int nsteps_run, int32_t *steps, size_t capacity,
                          size_t *count);
/* Ez.txt (one "%12f " line of Ez down each sieved column per snapshot), the
 * newline-only Ex/Hx/Hy/Hz.txt, the empty Ey.txt, and on close
 * FieldSaveDat.txt ("%d %d": lines per snapshot, snapshots). ez is laid out
 * as fdtd2d_read_field returns it. */
typedef struct fdtd2d_snapshots fdtd2d_snapshots;
int fdtd2d_snapshots_open(const fdtd2d_setup *s, const char *dir, fdtd2d_snapshots **out);
int fdtd2d_snapshots_add(fdtd2d_snapshots *w, const float *ez, size_t count);
int fdtd2d_snapshots_close(fdtd2d_snapshots *w, int *saved);

/* ------------------------------------------------------------------ plans
 * Replaces DevMemoryAllocation / the DevFieldCalculation loop /
 * DevMemoryFree (F/kernel.cu:1087-1618, 2065-2309, 2820-2987). */

/* Copies the problem to device `device`; all state starts at zero. */
int fdtd2d_plan_create(const fdtd2d_problem *p, int device, fdtd2d_plan **out);
/* As fdtd2d_plan_create, allocating nothing on the device and creating no
 * stream, event or graph. No counterpart in the reference: for callers that
 * cannot call hipMalloc (code under stream capture, serving loops that carve
 * shots from an arena of their own, harnesses that forbid allocation).
 *   mem, bytes: device memory on `device`, 256-byte aligned, at least
 *     fdtd2d_plan_footprint(p) bytes, owned by the caller and used by the
 *     plan until fdtd2d_plan_destroy (zeroed here, on `stream`, by a kernel);
 *   stream: the hipStream_t all of the plan's work goes to (NULL: the null
 *     stream); the plan's second kernel, where the grid needs one, runs after
 *     the first on that stream instead of beside it on a stream of its own.
 * The step kernels' tile lists live in a static 1 MiB device array of the
 * library (65536 entries of 16 bytes), one per device, held until the
 * process ends: a launch shape needing more is FDTD2D_EINVAL (here or at
 * fdtd2d_plan_set_option), and since every such plan on a device would share
 * that array, only one may exist per device at a time (FDTD2D_ESTATE);
 * plans of fdtd2d_plan_create are unaffected. Everything else (run, read,
 * write, options, destroy, which frees nothing) works as for any plan. */
int fdtd2d_plan_create_in(const fdtd2d_problem *p, int device, void *mem, size_t bytes, void *stream,
                          fdtd2d_plan **out);
/* Work on `stream` (a hipStream_t; NULL = the plan's own stream). */
int fdtd2d_plan_set_stream(fdtd2d_plan *plan, void *stream);
/* Time-blocking factor per launch (1..8; 0 = autotuned default). */
int fdtd2d_plan_set_tblock(fdtd2d_plan *plan, int steps_per_launch);
/* Launch-shape options: "tblock" (steps per launch, 1..8), "band" (rows per
   tile, >= tblock; 0 = automatic), "band_cpml" (rows per tile of the tiles
   that touch the CPML, 0 = automatic), "cols_per_lane" (1, 2 or 4), "prefetch"
   (rows loaded ahead per wave: 1, 2 or 4 with 1 column per lane, 1 or 2 with
   2, 1 with 4), "jz_first" (1, the default: interior work-groups that update
   Jz are launched first on each XCD; 0: column order only), "fused" (1, the
   default: one kernel per launch runs the interior tiles and every
   CPML-class tile but a few kinds - the general ones, and for the table
   medium the corners - which run in a second kernel beside it only where
   the grid has them; 0: all CPML-class tiles in the second kernel; with 4
   columns per lane launches are never fused). Results are identical for
   every choice. */
int fdtd2d_plan_set_option(fdtd2d_plan *plan, const char *name, double value);
/* Plan facts by name: "pitch", "tblock", "band", "band_cpml", "ntx", "nty",
   "workgroups", "workgroups_cpml" (work-groups of the second kernel),
   "tile_cols", "bytes", "steps", "cols_per_lane", "prefetch", "jz_first",
   "fused" (1 when launches are fused), "cells_interior", "cells_cpml",
   "tiles_interior", "tiles_cpml", "tiles_nojz" (interior tiles with no Jz
   update), "tiles_plasma_full" (interior tiles whose window has Ne > 0 in
   every cell), "cells_xstrip", "cells_ystrip", "cells_corner",
   "cells_events", "cells_cpml_general" (the CPML-class cells by the kind of
   tile that owns them: x-strip tiles carry only the x-strip psi and column
   tables, y-strip tiles only the y-strip psi and row records, corner tiles
   both, event tiles only the sources, detectors and captures, general tiles
   all of these), "cus" (compute units of the plan's device). */
int fdtd2d_plan_get_info(const fdtd2d_plan *plan, const char *name, double *value);
/* Advance nsteps steps (asynchronous on the plan's stream). */
int fdtd2d_run(fdtd2d_plan *plan, int nsteps);
int fdtd2d_sync(fdtd2d_plan *plan);
int fdtd2d_steps_done(const fdtd2d_plan *plan, int *n);
/* Traces: det [(nsteps+1)*ndet], comp [3][(nsteps+1)*ncomp] as Ez, Hx, Hy. */
int fdtd2d_read_detectors(fdtd2d_plan *plan, float *det, size_t det_count, float *comp,
                          size_t comp_count);
int fdtd2d_read_field(fdtd2d_plan *plan, int which, float *out, size_t count);
int fdtd2d_write_field(fdtd2d_plan *plan, int which, const float *in, size_t count);
/* psi strips, reference layouts (F/kernel.cu:1172-1214): HXY [2][npml-1][nx],
 * HYX [2][ny][npml-1], EZX [2][ny][npml], EZY [2][npml][nx]. */
int fdtd2d_read_psi(fdtd2d_plan *plan, int which, float *out, size_t count);
int fdtd2d_write_psi(fdtd2d_plan *plan, int which, const float *in, size_t count);
/* The tile partition a plan of this shape launches, computed on the host (no
   device needed), for uniform den tables and the nev event cells (sources,
   detectors, probes) at (ev_i[e], ev_j[e]): entries[5k..5k+4] = {first
   column, end column, first row, end row, class} of the cells wave k stores,
   for k < *count; class 0 interior kernel (cols_per_lane columns per lane),
   1-5 CPML-class tiles (always one column per lane): 1 general, 2 x-strip,
   3 y-strip, 4 corner, 5 event tiles. A CPML-class tile of the interior geometry is split into
   pieces of 64 - 2*tblock columns; the interior tiles start after, and end
   before, margins of such pieces as wide as the x-strip and the halo need,
   whose pieces at the grid's edges store up to 64 - tblock columns (no halo
   on the edge side).
   geometry (4 ints, may be null) = {vw, halo columns, rows per band, rows per
   CPML sub-band} of the interior tiles. entries may be null to query *count.
   tblock 0 = the default. No reference counterpart: the reference launches
   one thread per cell (F/kernel.cu:2027-2061). */
int fdtd2d_tile_lists(int nx, int ny, int npml, int tblock, int cols_per_lane, int band, int band_cpml,
                      const int *ev_i, const int *ev_j, int nev, int *entries, size_t capacity, size_t *count,
                      int *geometry);
/* RC tracker traces, [(nsteps+1) * nrc] (row n = step n, row 0 = 0). */
int fdtd2d_read_rc(fdtd2d_plan *plan, int32_t *spd, size_t count);
/* Processing line captures, [rows * line_size] with rows =
 * max(0, min(line_end, nsteps) - line_begin); row t = step line_begin + t + 1
 * (the reference's proc_field, F/FDTD_cuda.cu:471-483). Rows of steps not
 * yet run hold 0. */
int fdtd2d_read_line(fdtd2d_plan *plan, float *line, size_t count);
/* Device bytes this plan holds (the footprint admission control counts). */
int fdtd2d_plan_footprint(const fdtd2d_problem *p, uint64_t *bytes);
int fdtd2d_plan_destroy(fdtd2d_plan *plan);

/* Average duration (ms) of the interior step-kernel launches (the kernel that
 * updates the cells of the interior work-groups, `cells_interior` of
 * fdtd2d_plan_get_info, for up to `tblock` steps per launch) since timing was
 * enabled, from HIP events recorded around each launch on the plan's stream;
 * 0 disables timing. */
int fdtd2d_plan_time_kernels(fdtd2d_plan *plan, int enable);
int fdtd2d_plan_kernel_ms(fdtd2d_plan *plan, double *avg_ms, int *count);
/* The same for either kernel class: cls 0 the interior kernel (as
 * fdtd2d_plan_kernel_ms), 1 the CPML-class kernel (the work-groups whose
 * windows touch a CPML strip, the outer ring or an event cell, `cells_cpml`
 * of fdtd2d_plan_get_info; timed on the plan's second stream, where it runs
 * beside the interior kernel). count = launches timed. */
int fdtd2d_plan_kernel_time(fdtd2d_plan *plan, int cls, double *avg_ms, int *count);

/* ------------------------------------------------------------------ pools
 * Many independent shots (any grid size, medium, sources, detectors and step
 * count; one medium kind per pool) advanced together by persistent kernels
 * over one packed tile queue: the batched and served forms of the stepper.
 * No reference counterpart: the reference runs one grid per process
 * (F/FDTD_cuda.cu:384-441). Results of every shot are bitwise those of a
 * single plan of the same problem.
 *
 * A pool launch advances every unfinished shot by min(tblock, steps left)
 * (with time slicing, a sliced shot only at its turns);
 * fdtd2d_pool_step queues `nlaunch` of them (asynchronously; chunks of
 * `graph_chunk` launches are replayed from a captured hipGraph when no shot
 * ends inside them). When only sliced shots run, the launches between their
 * turns are skipped: they launch nothing but count in "launches" and in
 * fdtd2d_pool_wait's max_pending. Shots can be added and removed between
 * steps.
 *
 * Options (fdtd2d_pool_set_option): "tblock" (1..8, default 8), "band"
 * (default 128, 0 = per shot), "band_cpml" (default 64, 0 = per shot),
 * "prefetch" (1, the default, or 2), "cols_per_lane" (1, the default, or 2:
 * columns per lane of the interior tiles), "fused" (0, the default: one
 * persistent kernel per tile class on two streams; 1: one kernel per launch
 * that takes the CPML-class tiles first, then the interior ones, with the
 * CPML-class kinds it does not carry in a second kernel beside it), "wg_per_cu" /
 * "wg_per_cu_cpml" (persistent work-groups per CU of the interior or fused
 * kernel / of the CPML-class kernel, default 3 / 2), "graph_chunk" (0..16,
 * 0 = no graphs), "list_buffers" (2..4, default 4: device copies of the
 * tile lists, each kept for the shot set it holds; a rebuild of the lists
 * waits for the launches that read the copy it reuses), "slice_cells" and
 * "slice_every" (time slicing, default off: shots of at least slice_cells
 * cells run only in every slice_every-th launch, so that small shots
 * finish without waiting for a large shot's full launches; graphs are not
 * used while such a shot runs), "band_small" / "band_cpml_small" (band and
 * CPML sub-band rows of the other shots while slicing is on, 0 = "band" /
 * "band_cpml": a launch carrying only small shots lasts about one band's
 * row loop),
 * "budget_bytes" (before any shot: carve shots from one arena of that size
 * and refuse shots that do not fit, FDTD2D_ENOMEM).
 * Info (fdtd2d_pool_get_info): the options, "bytes_in_use", "largest_free",
 * "shots", "running", "launches", "cell_updates", "graph_replays",
 * "eager_launches", "tail_launches", "cus", and host-side counters:
 * "rebuilds" / "captures" (main tile lists rebuilt, graphs captured),
 * "list_switches" (launches that took a kept list copy instead),
 * "rebuild_s", "capture_s", "add_s" (host seconds spent in them and in
 * fdtd2d_pool_add), "blocked_s" (host seconds waiting for launches still on
 * the device, outside fdtd2d_pool_wait and fdtd2d_pool_sync). */
typedef struct fdtd2d_pool fdtd2d_pool;
int fdtd2d_pool_create(int device, fdtd2d_pool **out);
int fdtd2d_pool_set_option(fdtd2d_pool *pool, const char *name, double value);
int fdtd2d_pool_get_info(const fdtd2d_pool *pool, const char *name, double *value);
/* Copies the problem into the pool; *shot = its id (ids of removed shots are reused). */
int fdtd2d_pool_add(fdtd2d_pool *pool, const fdtd2d_problem *p, int *shot);
int fdtd2d_pool_step(fdtd2d_pool *pool, int nlaunch);
int fdtd2d_pool_sync(fdtd2d_pool *pool);
/* Blocks until at most max_pending of the launches issued so far are still
   unfinished on the device (keeps a serving loop a bounded distance ahead). */
int fdtd2d_pool_wait(fdtd2d_pool *pool, int max_pending);
/* complete = all steps issued and finished on the device (does not block). */
int fdtd2d_pool_shot_state(fdtd2d_pool *pool, int shot, int *steps_done, int *steps_total, int *complete);
/* As fdtd2d_read_detectors / fdtd2d_read_field for one shot; waits only for
   the launches the shot's results depend on. */
int fdtd2d_pool_read_detectors(fdtd2d_pool *pool, int shot, float *det, size_t det_count, float *comp,
                               size_t comp_count);
int fdtd2d_pool_read_field(fdtd2d_pool *pool, int shot, int which, float *out, size_t count);
int fdtd2d_pool_read_line(fdtd2d_pool *pool, int shot, float *line, size_t count);
int fdtd2d_pool_remove(fdtd2d_pool *pool, int shot);
int fdtd2d_pool_destroy(fdtd2d_pool *pool);

#ifdef __cplusplus
}
#endif
#endif /* FDTD2D_H */
