// One RK4 stage of the stage-fused shallow-water step, shared by the fused
// driver and the adjoint's recording forward: the buffers each stage reads
// and writes are named explicitly, so a caller can point them at a tape.
#pragma once
#include "ctx.h"

namespace mpasx {

struct StageIO {
  int stage, scheme, freeze_u, use_visc, use_hmix;
  // 1: the tracer-only (ocean) run: u, h and h_edge fixed, the FCT phases
  // only, ocean tracer bookkeeping
  int ocean;
  // 1: the limiter factors in `rfac` are given (a frozen-limiter replay)
  // and k_stage_cells leaves them untouched
  int keep_rfac;
  // levels per lane and level chunks per entity of the levels-on-lanes
  // kernels (ensembles); unused by the single-run kernels
  int lpl, nchunk;
  double w_rk, w_sub, dt, visc_coef, kappa;
  // state read by the stage (the reference's provis) and its diagnostics
  const double *xu, *xh, *xtr;
  const double *h_edge, *ke, *pv_edge, *vort, *div;
  const double *h_s;
  // state at the start of the step (the reference's prev)
  const double *u0, *h0, *tr0;
  // RK4 accumulators (the reference's next; tracers h-weighted)
  double *nu, *nh, *ntr;
  // state written by the stage: next provis, or the final state at stage 3
  double *yu, *yh, *ytr;
  double *ke_y, *div_y;
  // FCT intermediates
  double *anti, *rfac, *tlo;
};

// RK4 weights of apps/sw/run.cc:23-24.
inline double rk_weight(int s, double dt) {
  const double w[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  return w[s];
}
inline double rk_substep_weight(int s, double dt) {
  const double w[4] = {dt / 2, dt / 2, dt, 0};
  return w[s];
}

// Stage parameters (everything but the buffers) from the context.
StageIO stage_params(const mpas_ctx *ctx, int s, double dt);

// Launches one stage and the diagnostics of its output (vorticity and
// pv_vertex, then h_edge and pv_edge).
int launch_stage(mpas_ctx *ctx, const StageIO &S, double *vort_y, double *pv_vertex,
                 double *h_edge_y, double *pv_edge_y);

// The same on an explicit mesh and stream (no context, no profiling).
int launch_stage_dm(const DMesh &M, hipStream_t st, const StageIO &S, double *vort_y,
                    double *pv_vertex, double *h_edge_y, double *pv_edge_y);

// compute_solve_diagnostics of (u, h) into explicit buffers (ke and the
// divergence by cells, vorticity and pv by vertices, h_edge and pv_edge).
int diagnostics_into(mpas_ctx *ctx, const double *u, const double *h, double *ke, double *div,
                     double *vort, double *pv_vertex, double *h_edge, double *pv_edge);

int diagnostics_dm(const DMesh &M, hipStream_t st, int use_visc, const double *u,
                   const double *h, double *ke, double *div, double *vort, double *pv_vertex,
                   double *h_edge, double *pv_edge);

// diagnostics_into the context's ping-pong slot `slot`.
int sw_diagnostics(mpas_ctx *ctx, const double *u, const double *h, int slot);

// nsteps RK4 steps of the tracer-only (ocean) run on the resident state:
// tracers advected by the fixed u with the fixed h and h_edge.
int ocean_rk4_steps(mpas_ctx *ctx, double dt, int nsteps);

// nsteps fused RK4 steps of the resident state.  frozen_rfac, when given,
// holds per (step, stage) the limiter factors to use instead of computing
// them (frozen_rfac[4 * step + stage]).
int sw_rk4_steps(mpas_ctx *ctx, double dt, int nsteps, double *const *frozen_rfac = nullptr);

// The buffer roles of stage s of the fused step over the buffers of W.
StageIO sw_stage_io(const mpas_ctx *ctx, const SwState &W, int s, double dt);

// nsteps fused RK4 steps over the buffers of W (not the context's state),
// launched on the context's stream without profiling.
int sw_rk4_steps_on(const mpas_ctx *ctx, SwState &W, double dt, int nsteps);

}  // namespace mpasx
