// Initial states on generated meshes (initial.cc): the reference's test
// cases and the seeded states of the workloads.
#pragma once
#include <stdint.h>

#include "meshgen.h"

#ifdef __cplusplus
extern "C" {
#endif

int mg_init_sw_balanced(const mg_mesh *m, double a, double omega, double gravity, double u0,
                        double h_mean, double bump_amp, double bump_radius, uint64_t seed,
                        double *u, double *h, double *centre_latlon);
int mg_init_tracer_blobs(const mg_mesh *m, double a, int nblobs, double rmin, double rmax,
                         uint64_t seed, double *tr);
int mg_init_sw_case1(const mg_mesh *m, double a, double *u, double *h);
int mg_init_sw_case5(const mg_mesh *m, double a, double omega, double gravity, double *u,
                     double *h, double *h_s, double *fVertex, double *tr, int ntr);
int mg_init_sw_case6(const mg_mesh *m, double a, double omega, double gravity, double *u,
                     double *h);
int mg_init_ocean_case(const mg_mesh *m, int which, double a, double coef, double elapsed,
                       double *normalVelocity, double *layerThickness, double *tr);
/* The end-of-run error report of the ocean test cases (calc_error_case_1/2/3,
   apps/ocean/tasks/testcase.cc:175-294) for tr[c * stride] (tracer 0 of
   level 0): relative errors against the case's solution after `elapsed`
   seconds, l2 = sqrt(sum of squares) / nCells, and max_err = |errMax|
   with the reference's update rule (errMax = error whenever |error| >
   errMax, so one negative error makes it follow later cells). */
int mg_ocean_error(const mg_mesh *m, int which, double a, double coef, double elapsed,
                   const double *tr, int64_t stride, double *l2, double *max_err);

#ifdef __cplusplus
}
#endif
