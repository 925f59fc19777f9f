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

#ifdef __cplusplus
}
#endif
