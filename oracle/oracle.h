/*
 * oracle.h - CPU restatement of the reference's TMz update equations.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP stepper is
 * compared against; only tests, the smoke check and the CPU-baseline leg of
 * the benchmark may load it. The product never calls it.
 *
 * Single-threaded by default (optional row-parallel threads give bit-identical
 * results: every phase is a Jacobi sweep with no reductions). Precision 32
 * reproduces the reference's FP32 arithmetic, including the FP64 promotion of
 * the Jz update (F/kernel.cu:127, 155, 2102); precision 64 runs the same
 * loops in double from the same float tables.
 */
#ifndef FDTD2D_ORACLE_H
#define FDTD2D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/fdtd2d.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_sim oracle_sim;

int oracle_create(const fdtd2d_problem *p, int precision, oracle_sim **out);
int oracle_set_threads(oracle_sim *s, int nthreads);
int oracle_run(oracle_sim *s, int nsteps);
int oracle_steps_done(const oracle_sim *s);
/* Fields in the reference layouts (see fdtd2d.h), as double. */
int oracle_read_field(const oracle_sim *s, int which, double *out, size_t count);
int oracle_write_field(oracle_sim *s, int which, const double *in, size_t count);
int oracle_read_psi(const oracle_sim *s, int which, double *out, size_t count);
int oracle_write_psi(oracle_sim *s, int which, const double *in, size_t count);
/* det [(nsteps+1)*ndet], comp [3][(nsteps+1)*ncomp] (Ez, Hx, Hy). */
int oracle_read_detectors(const oracle_sim *s, double *det, size_t det_count, double *comp,
                          size_t comp_count);
/* RC tracker traces [(nsteps+1)*nrc], row n = step n. */
int oracle_read_rc(const oracle_sim *s, int32_t *spd, size_t count);
/* Processing line captures [rows][line_size], row t = step line_begin+t+1. */
int oracle_read_line(const oracle_sim *s, double *line, size_t count);
/* processing_core's time-of-flight search (oracle/processing.cpp), with the
 * arguments and results of fdtd2d_processing_tof. */
int oracle_processing_tof(const fdtd2d_processing *a, const float *line, int times, const float *weights,
                          int pass, float *sd, float *cf, float *arc, float *original, float *filtered);
void oracle_destroy(oracle_sim *s);
const char *oracle_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
