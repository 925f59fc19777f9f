// One kernel (or one kernel per phase) for each reference task.  Where the
// reference scatters (`x(cell1) -= ...; x(cell2) += ...` over all edges in
// ascending id) the kernels gather over the same contributions in the same
// ascending order, so every sum is formed operation for operation as the
// reference forms it and no atomics are needed.  Every expression keeps the
// reference's association; the file is compiled with -ffp-contract=off.
#include <algorithm>
#include <cmath>

#include "ctx.h"
#include "ops.h"

namespace mpasx {
namespace {

constexpr int BS = 256;

#define IDX_T(c, k, t) ((size_t(c) * M.nL + (k)) * M.nT + (t))

// ---- apps/sw/tasks/solve.cc:11-107 (compute_solve_diagnostics) ----------

// circulation, vorticity, h_vertex, pv_vertex (solve.cc:42-60, 81-92)
__global__ __launch_bounds__(BS) void k_diag_vertices(DMesh M, const double *__restrict__ h,
                                                      const double *__restrict__ u,
                                                      double *h_vertex, double *circulation,
                                                      double *vorticity, double *pv_vertex) {
  const int v = gtid();
  if (v < M.nV) diag_vertex(M, v, h, u, h_vertex, circulation, vorticity, pv_vertex);
}

// ke (solve.cc:63-77), edges of the cell in file order
__global__ __launch_bounds__(BS) void k_diag_ke(DMesh M, const double *__restrict__ u,
                                                double *ke) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int n = M.nEoC[c];
  const double ar = M.area[c];
  for (int k = 0; k < M.nL; k++) {
    double s = 0.0;
    for (int j = 0; j < n; j++) {
      const int e = M.c2e[j * M.nC + c];
      const double ue = u[size_t(e) * M.nL + k];
      s += 0.25 * M.dc[e] * M.dv[e] * (ue * ue);
    }
    ke[size_t(c) * M.nL + k] = s / ar;
  }
}

// h_edge (solve.cc:31-39) and pv_edge (solve.cc:95-106)
__global__ __launch_bounds__(BS) void k_diag_edges(DMesh M, const double *__restrict__ h,
                                                   const double *__restrict__ pv_vertex,
                                                   double *h_edge, double *pv_edge) {
  const int e = gtid();
  if (e < M.nE) diag_edge(M, e, h, pv_vertex, h_edge, pv_edge);
}

// ---- solve.cc:169-234 (compute_tend) --------------------------------------

// tend_h: per cell, its edges in ascending id (the scatter order, 191-201)
__global__ __launch_bounds__(BS) void k_tend_h(DMesh M, const double *__restrict__ u,
                                               const double *__restrict__ h_edge,
                                               double *tend_h) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int n = M.nEoC[c];
  const uint32_t perm = M.ascPerm[c];
  const uint8_t neg = M.sgnNeg[c];
  for (int k = 0; k < M.nL; k++) {
    double t = 0;
    for (int r = 0; r < n; r++) {
      const int j = (perm >> (4 * r)) & 15;
      const int e = M.c2e[j * M.nC + c];
      const double flux = u[size_t(e) * M.nL + k] * M.dv[e] * h_edge[size_t(e) * M.nL + k];
      if ((neg >> j) & 1)
        t -= flux;
      else
        t += flux;
    }
    tend_h[size_t(c) * M.nL + k] = t / M.area[c];
  }
}

// tend_u (solve.cc:215-233): PV flux through weightsOnEdge, KE and height
// gradients
__global__ __launch_bounds__(BS) void k_tend_u(DMesh M, const double *__restrict__ u,
                                               const double *__restrict__ h_edge,
                                               const double *__restrict__ pv_edge,
                                               const double *__restrict__ ke,
                                               const double *__restrict__ h,
                                               const double *__restrict__ h_s, double *tend_u) {
  const int e = gtid();
  if (e >= M.nE) return;
  const int2 cc = M.e2c[e];
  for (int k = 0; k < M.nL; k++)
    tend_u[size_t(e) * M.nL + k] = tend_u_at(M, e, k, cc, u, h_edge, pv_edge, ke, h, h_s);
}

// ---- solve.cc:237-270 (compute_scalar_tend) -------------------------------
__global__ __launch_bounds__(BS) void k_scalar_tend(DMesh M, const double *__restrict__ h_edge,
                                                    const double *__restrict__ u,
                                                    const double *__restrict__ tr,
                                                    double *tend) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int n = M.nEoC[c];
  const uint32_t perm = M.ascPerm[c];
  const uint8_t neg = M.sgnNeg[c];
  const double ar = M.area[c];
  for (int k = 0; k < M.nL; k++)
    for (int t = 0; t < M.nT; t++) {
      double s = 0.0;
      for (int r = 0; r < n; r++) {
        const int j = (perm >> (4 * r)) & 15;
        const int e = M.c2e[j * M.nC + c];
        const int2 cc = M.e2c[e];
        const double tracer_edge = 0.5 * (tr[IDX_T(cc.x, k, t)] + tr[IDX_T(cc.y, k, t)]);
        const double flux =
            u[size_t(e) * M.nL + k] * M.dv[e] * h_edge[size_t(e) * M.nL + k] * tracer_edge;
        if ((neg >> j) & 1)
          s -= flux / ar;
        else
          s += flux / ar;
      }
      tend[IDX_T(c, k, t)] = s;
    }
}

// ---- RK4 bookkeeping (solve.cc:110-166, 273-401) ---------------------------
__global__ __launch_bounds__(BS) void k_axpy_edges(int n, double *out, const double *a,
                                                   const double *t, double w, int mode) {
  const int i = gtid();
  if (i >= n) return;
  if (mode == 0) out[i] = a[i] + w * t[i];  // compute_substep
  else out[i] += w * t[i];                  // accumulate_update
}

// compute_substep, cells (solve.cc:291-305)
__global__ __launch_bounds__(BS) void k_substep_cells(DMesh M, double *h_p, const double *h_old,
                                                      const double *h_tend, double *tr_p,
                                                      const double *tr_old,
                                                      const double *tr_tend, double w) {
  const int c = gtid();
  if (c >= M.nC) return;
  for (int k = 0; k < M.nL; k++) {
    const size_t o = size_t(c) * M.nL + k;
    const double hp = h_old[o] + w * h_tend[o];
    h_p[o] = hp;
    for (int t = 0; t < M.nT; t++)
      tr_p[IDX_T(c, k, t)] = (h_old[o] * tr_old[IDX_T(c, k, t)] + w * tr_tend[IDX_T(c, k, t)]) / hp;
  }
}

// accumulate_update, cells (solve.cc:332-340)
__global__ __launch_bounds__(BS) void k_accumulate_cells(DMesh M, double *h_new,
                                                         const double *h_tend, double *tr_new,
                                                         const double *tr_tend, double w) {
  const int c = gtid();
  if (c >= M.nC) return;
  for (int k = 0; k < M.nL; k++) {
    const size_t o = size_t(c) * M.nL + k;
    h_new[o] += w * h_tend[o];
    for (int t = 0; t < M.nT; t++) tr_new[IDX_T(c, k, t)] += w * tr_tend[IDX_T(c, k, t)];
  }
}

// setup_timestep, cells (solve.cc:158-165): tracers become h-weighted
__global__ __launch_bounds__(BS) void k_setup_cells(DMesh M, const double *h_old,
                                                    const double *tr_old, double *h_new,
                                                    double *tr_new) {
  const int c = gtid();
  if (c >= M.nC) return;
  for (int k = 0; k < M.nL; k++) {
    const size_t o = size_t(c) * M.nL + k;
    h_new[o] = h_old[o];
    for (int t = 0; t < M.nT; t++) tr_new[IDX_T(c, k, t)] = tr_old[IDX_T(c, k, t)] * h_old[o];
  }
}

// finalize_timestep, cells (solve.cc:351-357)
__global__ __launch_bounds__(BS) void k_finalize_cells(DMesh M, const double *h_new,
                                                       double *tr_new) {
  const int c = gtid();
  if (c >= M.nC) return;
  for (int k = 0; k < M.nL; k++)
    for (int t = 0; t < M.nT; t++) tr_new[IDX_T(c, k, t)] /= h_new[size_t(c) * M.nL + k];
}

// ---- del2 viscosity on u (no reference counterpart) ------------------------
// divergence(c) = sum over edges (file order) of outward dvEdge * u / areaCell;
// outward = -edgeSignOnCell.
__global__ __launch_bounds__(BS) void k_divergence(DMesh M, const double *__restrict__ u,
                                                   double *div) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int n = M.nEoC[c];
  const uint8_t neg = M.sgnNeg[c];
  for (int k = 0; k < M.nL; k++) {
    double s = 0.0;
    for (int j = 0; j < n; j++) {
      const int e = M.c2e[j * M.nC + c];
      s += (-sign_on_cell(neg, j)) * M.dv[e] * u[size_t(e) * M.nL + k];
    }
    div[size_t(c) * M.nL + k] = s / M.area[c];
  }
}

__global__ __launch_bounds__(BS) void k_visc(DMesh M, const double *__restrict__ div,
                                             const double *__restrict__ vort, double coef,
                                             double *tend_u) {
  const int e = gtid();
  if (e >= M.nE) return;
  const int2 cc = M.e2c[e], vv = M.e2v[e];
  for (int k = 0; k < M.nL; k++)
    tend_u[size_t(e) * M.nL + k] += visc_at(M, e, k, cc, vv, div, vort, coef);
}

// ---- apps/ocean/tasks/solve.cc ----------------------------------------------

// calc_normal_thickness_flux (40-57)
__global__ __launch_bounds__(BS) void k_ntf(int n, const double *lte, const double *nv,
                                            double *ntf) {
  const int i = gtid();
  if (i < n) ntf[i] = nv[i] * lte[i];
}

// calc_provisional_layer_thicknesses (60-97), edges in file order
__global__ __launch_bounds__(BS) void k_hprov(DMesh M, const double *__restrict__ lt,
                                              const double *__restrict__ ntf, double *hProv,
                                              double *hProvInv, double *hNewInv, double dt) {
  const int c = gtid();
  if (c >= M.nC) return;
  const double dtInvAreaCell = dt / M.area[c];
  const int kmax = lvl_cell(M, c);
  const int n = M.nEoC[c];
  const uint8_t neg = M.sgnNeg[c];
  for (int k = 0; k < kmax; k++) {
    double hp = lt[size_t(c) * M.nL + k];
    for (int j = 0; j < n; j++) {
      const int e = M.c2e[j * M.nC + c];
      const double signedFactor = dtInvAreaCell * M.dv[e] * sign_on_cell(neg, j);
      hp += signedFactor * ntf[size_t(e) * M.nL + k];
    }
    hProv[size_t(c) * M.nL + k] = hp;
    hProvInv[size_t(c) * M.nL + k] = 1.0 / hp;
    hNewInv[size_t(c) * M.nL + k] = 1.0 / (hp);
  }
}

// calc_horiz_tracer_advection_flux_mono_low_order (99-118)
__global__ __launch_bounds__(BS) void k_low_order(DMesh M, const double *__restrict__ ntf,
                                                  const double *__restrict__ tr, double *lo) {
  const int e = gtid();
  if (e >= M.nE) return;
  const int2 cc = M.e2c[e];
  for (int k = 0; k < M.nL; k++) {
    const double f = ntf[size_t(e) * M.nL + k];
    const double ntf1 = smax(0.0, f), ntf2 = smin(0.0, f);
    for (int t = 0; t < M.nT; t++)
      lo[(size_t(e) * M.nL + k) * M.nT + t] =
          M.dv[e] * (ntf1 * tr[IDX_T(cc.x, k, t)] + ntf2 * tr[IDX_T(cc.y, k, t)]);
  }
}

// calc_horiz_tracer_advection_flux_mono_high_order (120-167); each (k, t)
// sum runs over the advection cells in list order, as flxTmp[k][t] does.
__global__ __launch_bounds__(BS) void k_high_order(DMesh M, const double *__restrict__ ntf,
                                                   const double *__restrict__ tr, double *ho) {
  const int e = gtid();
  if (e >= M.nE) return;
  const int n = M.nAdv[e];
  for (int k = 0; k < M.nL; k++) {
    const double f = ntf[size_t(e) * M.nL + k];
    const double mask = M.hoMask ? double(M.hoMask[size_t(e) * M.nL + k]) : 1.0;
    const double wgt = f * mask;
    const double sgn = (f < 0.0) ? -1.0 : double(f > 0.);
    for (int t = 0; t < M.nT; t++) {
      double s = 0.0;
      for (int i = 0; i < n; i++) {
        const int ic = M.adv[size_t(i) * M.nE + e];
        if (k >= lvl_cell(M, ic)) continue;
        const double coef1 = M.advCoef[size_t(i) * M.nE + e];
        const double coef3 = M.advCoef3[size_t(i) * M.nE + e] * M.coef3rd;
        const double weightFactor = wgt * (coef1 + coef3 * sgn);
        s += tr[IDX_T(ic, k, t)] * weightFactor;
      }
      ho[(size_t(e) * M.nL + k) * M.nT + t] = s;
    }
  }
}

// correct_and_rescale_high_order_mono_horiz_flux (169-285), phase 2
// (216-229): antidiffusive flux
__global__ __launch_bounds__(BS) void k_fct_correct(DMesh M, const double *__restrict__ ntf,
                                                    const double *__restrict__ lo,
                                                    const double *__restrict__ tr, double *ho) {
  const int e = gtid();
  if (e >= M.nE) return;
  const int2 cc = M.e2c[e];
  const int kmax = lvl_edge(M, e);
  for (int k = 0; k < kmax; k++) {
    const int mask = M.hoMask ? M.hoMask[size_t(e) * M.nL + k] : 1;
    const double tracerWeight =
        double((mask + 1) & 1) * (M.dv[e] * 0.5) * ntf[size_t(e) * M.nL + k];
    for (int t = 0; t < M.nT; t++) {
      const size_t o = (size_t(e) * M.nL + k) * M.nT + t;
      double x = ho[o];
      x += tracerWeight * (tr[IDX_T(cc.x, k, t)] + tr[IDX_T(cc.y, k, t)]);
      x -= lo[o];
      ho[o] = x;
    }
  }
}

// phases 1 and 3 (195-214, 232-267): bounds over cellsOnCell, upwind
// tendency, flux sums and the limiter factors, edges in file order
__global__ __launch_bounds__(BS) void k_fct_cell_limit(
    DMesh M, const double *__restrict__ hProvInv, const double *__restrict__ lt,
    const double *__restrict__ lo, const double *__restrict__ tr,
    const double *__restrict__ ho, double *flxIn, double *flxOut, double *tend,
    double *tracerMax, double *tracerMin, double dt) {
  const int c = gtid();
  if (c >= M.nC) return;
  const double eps = 1.e-10;
  const int kc = lvl_cell(M, c);
  const int n = M.nEoC[c];
  const uint8_t neg = M.sgnNeg[c];
  const double invAreaCell1 = 1.0 / M.area[c];
  for (int k = 0; k < M.nL; k++)
    for (int t = 0; t < M.nT; t++) {
      const size_t o = IDX_T(c, k, t);
      double tmax = 0, tmin = 0, fin, fout;
      if (k < kc) {
        tmax = tr[o];
        tmin = tr[o];
        for (int j = 0; j < n; j++) {
          const int c2 = M.c2c[j * M.nC + c];
          if (k < min(kc, lvl_cell(M, c2))) {
            tmax = smax(tmax, tr[IDX_T(c2, k, t)]);
            tmin = smin(tmin, tr[IDX_T(c2, k, t)]);
          }
        }
        fin = 0.0;
        fout = 0.0;
      } else {
        fin = flxIn[o];
        fout = flxOut[o];
      }
      double te = tend[o];
      for (int j = 0; j < n; j++) {
        const int e = M.c2e[j * M.nC + c];
        if (k >= lvl_edge(M, e)) continue;
        const double signedFactor = sign_on_cell(neg, j) * invAreaCell1;
        const size_t oe = (size_t(e) * M.nL + k) * M.nT + t;
        te += signedFactor * lo[oe];
        fin += smax(0.0, signedFactor * ho[oe]);
        fout += smin(0.0, signedFactor * ho[oe]);
      }
      tend[o] = te;
      if (k < kc) {
        const size_t ol = size_t(c) * M.nL + k;
        const double tracerUpwindNew = (tr[o] * lt[ol] + dt * te) * hProvInv[ol];
        const double tracerMaxNew = tracerUpwindNew + dt * fin * hProvInv[ol];
        const double tracerMinNew = tracerUpwindNew + dt * fout * hProvInv[ol];
        double scaleFactor = (tmax - tracerUpwindNew) / (tracerMaxNew - tracerUpwindNew + eps);
        fin = smin(1.0, smax(0.0, scaleFactor));
        scaleFactor = (tracerUpwindNew - tmin) / (tracerUpwindNew - tracerMinNew + eps);
        fout = smin(1.0, smax(0.0, scaleFactor));
        tracerMax[o] = tmax;
        tracerMin[o] = tmin;
      }
      flxIn[o] = fin;
      flxOut[o] = fout;
    }
}

// phase 4 (270-282): rescale the antidiffusive flux
__global__ __launch_bounds__(BS) void k_fct_rescale(DMesh M, const double *__restrict__ flxIn,
                                                    const double *__restrict__ flxOut,
                                                    double *ho) {
  const int e = gtid();
  if (e >= M.nE) return;
  const int2 cc = M.e2c[e];
  const int kmax = lvl_edge(M, e);
  for (int k = 0; k < kmax; k++)
    for (int t = 0; t < M.nT; t++) {
      const size_t o = (size_t(e) * M.nL + k) * M.nT + t;
      const double x = ho[o];
      ho[o] = smax(0.0, x) * smin(flxOut[IDX_T(cc.x, k, t)], flxIn[IDX_T(cc.y, k, t)]) +
              smin(0.0, x) * smin(flxIn[IDX_T(cc.x, k, t)], flxOut[IDX_T(cc.y, k, t)]);
    }
}

// accumulate_horiz_tracer_advection_tend_mono_FCT (287-326)
__global__ __launch_bounds__(BS) void k_fct_accumulate(DMesh M, const double *__restrict__ hProvInv,
                                                       const double *__restrict__ lt,
                                                       const double *__restrict__ ho,
                                                       const double *__restrict__ tr,
                                                       double *trProv, double *tend, double dt) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int n = M.nEoC[c];
  const uint8_t neg = M.sgnNeg[c];
  const double invAreaCell1 = 1.0 / M.area[c];
  const int kc = lvl_cell(M, c);
  for (int k = 0; k < M.nL; k++)
    for (int t = 0; t < M.nT; t++) {
      const size_t o = IDX_T(c, k, t);
      double te = tend[o];
      for (int j = 0; j < n; j++) {
        const int e = M.c2e[j * M.nC + c];
        if (k >= lvl_edge(M, e)) continue;
        const double signedFactor = sign_on_cell(neg, j) * invAreaCell1;
        te += signedFactor * ho[(size_t(e) * M.nL + k) * M.nT + t];
      }
      tend[o] = te;
      if (k < kc && trProv) {
        const size_t ol = size_t(c) * M.nL + k;
        trProv[o] = (tr[o] * lt[ol] + dt * te) * hProvInv[ol];
      }
    }
}

// tracer_hmix_del2_tend (328-366)
__global__ __launch_bounds__(BS) void k_hmix_del2(DMesh M, const double *__restrict__ lte,
                                                  const double *__restrict__ tr, double *tend,
                                                  double coef) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int n = M.nEoC[c];
  const uint8_t neg = M.sgnNeg[c];
  const double invArea = 1.0 / M.area[c];
  for (int k = 0; k < M.nL; k++)
    for (int t = 0; t < M.nT; t++) {
      const size_t o = IDX_T(c, k, t);
      double te = tend[o];
      for (int j = 0; j < n; j++) {
        const int e = M.c2e[j * M.nC + c];
        if (k >= lvl_edge(M, e)) continue;
        const int2 cc = M.e2c[e];
        const double rho_scaling = M.msd * coef * M.dv[e] / M.dc[e];
        const double tracer_turb_flux = tr[IDX_T(cc.y, k, t)] - tr[IDX_T(cc.x, k, t)];
        const double flux = lte[size_t(e) * M.nL + k] * tracer_turb_flux * rho_scaling;
        te -= sign_on_cell(neg, j) * flux * invArea;
      }
      tend[o] = te;
    }
}

// step_tracers_forward (368-399)
__global__ __launch_bounds__(BS) void k_step_tracers(int n, const double *tend, const double *tr,
                                                     double *trNew, double *trPrev, double w,
                                                     double ws) {
  const int i = gtid();
  if (i >= n) return;
  trNew[i] += w * tend[i];
  trPrev[i] = tr[i] + ws * tend[i];
}

// plain streaming copy, 16 bytes per lane per iteration
// One element per lane over a grid that covers the buffer (the shape of the
// stage kernels): the fastest plain copy measured on MI355X.
__global__ __launch_bounds__(BS) void k_copy(int64_t n, const double *__restrict__ src,
                                             double *__restrict__ dst) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(BS) void k_fill(int64_t n, double *p, double v) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// compute_error (solve.cc:404-436): per-cell terms, then a fixed-order
// reduction on the host (the reference sums cells in ascending id)
__global__ __launch_bounds__(BS) void k_error_terms(DMesh M, const double *__restrict__ h,
                                                    const double *__restrict__ lat,
                                                    const double *__restrict__ lon, double a,
                                                    double elapsed, double *num, double *den) {
  const int c = gtid();
  if (c >= M.nC) return;
  const double pi = 3.141592653589793;
  const double h0 = 1000.0;
  const double day_frac = elapsed / 86400 / 12;
  const double rot_ang = (3 * pi / 2.) + 2 * pi * day_frac;
  const double s1 = sin(0.5 * (lat[c] - 0.0)), s2 = sin(0.5 * (lon[c] - rot_ang));
  const double arg1 = sqrt(s1 * s1 + cos(0.0) * cos(lat[c]) * (s2 * s2));
  const double r = 2 * a * asin(arg1);
  const double h_true = (r < a / 3.0) ? (h0 / 2.0) * (1.0 + cos(pi * r * 3.0 / a)) : h0 / 2.0;
  const double d = h[size_t(c) * M.nL] - h_true;
  num[c] = d * d * cos(lat[c]) / M.area[c];
  den[c] = h_true * h_true * cos(lat[c]) / M.area[c];
}

}  // namespace

#define CTX_CHECK(ctx)                    \
  do {                                    \
    int _rc = check_ctx(ctx);             \
    if (_rc) return _rc;                  \
  } while (0)
#define NONNULL(p) \
  if (!(p)) MPAS_FAIL(MPAS_EINVAL, std::string(__func__) + ": " #p " is NULL")
#define LAUNCH(kern, n, ...)                                                  \
  do {                                                                        \
    if ((n) > 0)                                                              \
      hipLaunchKernelGGL(kern, dim3(blocks_for((n), BS)), dim3(BS), 0,        \
                         ctx->stream, __VA_ARGS__);                           \
    int _rc = after_launch(#kern);                                            \
    if (_rc) return _rc;                                                      \
  } while (0)

static int need_fct(const mpas_ctx *ctx, const char *who) {
  if (!ctx->host.fct)
    MPAS_FAIL(MPAS_ESTATE, std::string(who) +
                               ": the context was created without the FCT setup "
                               "(tracer_scheme FCT or kappa2 != 0 builds it)");
  return 0;
}

}  // namespace mpasx

using namespace mpasx;

extern "C" {

int mpas_sw_compute_solve_diagnostics(mpas_ctx *ctx, const double *h, const double *u,
                                      double *h_edge, double *h_vertex, double *circulation,
                                      double *vorticity, double *ke, double *pv_edge,
                                      double *pv_vertex) {
  CTX_CHECK(ctx);
  NONNULL(h); NONNULL(u); NONNULL(h_edge); NONNULL(h_vertex); NONNULL(circulation);
  NONNULL(vorticity); NONNULL(ke); NONNULL(pv_edge); NONNULL(pv_vertex);
  const DMesh &M = ctx->dm;
  LAUNCH(k_diag_vertices, M.nV, M, h, u, h_vertex, circulation, vorticity, pv_vertex);
  LAUNCH(k_diag_ke, M.nC, M, u, ke);
  LAUNCH(k_diag_edges, M.nE, M, h, pv_vertex, h_edge, pv_edge);
  return 0;
}

int mpas_sw_init_provis(mpas_ctx *ctx, const double *u_old, const double *h_old,
                        const double *tracers_old, const double *h_edge_old,
                        const double *ke_old, const double *pv_edge_old, double *u_provis,
                        double *h_provis, double *tracers_provis, double *h_edge_provis,
                        double *ke_provis, double *pv_edge_provis) {
  CTX_CHECK(ctx);
  const DMesh &M = ctx->dm;
  const size_t cb = sizeof(double) * M.nC * M.nL, eb = sizeof(double) * M.nE * M.nL;
  const size_t tb = cb * M.nT;
  struct { void *d; const void *s; size_t n; } cp[] = {
      {h_provis, h_old, cb}, {ke_provis, ke_old, cb}, {tracers_provis, tracers_old, tb},
      {u_provis, u_old, eb}, {h_edge_provis, h_edge_old, eb}, {pv_edge_provis, pv_edge_old, eb}};
  for (auto &x : cp) {
    if (!x.n) continue;
    if (!x.d || !x.s) MPAS_FAIL(MPAS_EINVAL, "mpas_sw_init_provis: NULL field");
    MPAS_HIP(hipMemcpyAsync(x.d, x.s, x.n, hipMemcpyDeviceToDevice, ctx->stream));
  }
  return 0;
}

int mpas_sw_setup_timestep(mpas_ctx *ctx, const double *u_old, const double *h_old,
                           const double *tracers_old, double *u_new, double *h_new,
                           double *tracers_new) {
  CTX_CHECK(ctx);
  NONNULL(u_old); NONNULL(h_old); NONNULL(u_new); NONNULL(h_new);
  const DMesh &M = ctx->dm;
  if (M.nT) { NONNULL(tracers_old); NONNULL(tracers_new); }
  MPAS_HIP(hipMemcpyAsync(u_new, u_old, sizeof(double) * M.nE * M.nL, hipMemcpyDeviceToDevice,
                          ctx->stream));
  LAUNCH(k_setup_cells, M.nC, M, h_old, tracers_old, h_new, tracers_new);
  return 0;
}

int mpas_sw_compute_tend(mpas_ctx *ctx, const double *h_edge, const double *h,
                         const double *h_s, const double *u, const double *ke,
                         const double *pv_edge, double *tend_h, double *tend_u) {
  CTX_CHECK(ctx);
  NONNULL(h_edge); NONNULL(h); NONNULL(h_s); NONNULL(u); NONNULL(ke); NONNULL(pv_edge);
  NONNULL(tend_h); NONNULL(tend_u);
  const DMesh &M = ctx->dm;
  LAUNCH(k_tend_h, M.nC, M, u, h_edge, tend_h);
  LAUNCH(k_tend_u, M.nE, M, u, h_edge, pv_edge, ke, h, h_s, tend_u);
  return 0;
}

int mpas_sw_compute_scalar_tend(mpas_ctx *ctx, const double *h_edge, const double *u,
                                const double *tracers, double *tracer_tend) {
  CTX_CHECK(ctx);
  const DMesh &M = ctx->dm;
  if (!M.nT) return 0;
  NONNULL(h_edge); NONNULL(u); NONNULL(tracers); NONNULL(tracer_tend);
  LAUNCH(k_scalar_tend, M.nC, M, h_edge, u, tracers, tracer_tend);
  return 0;
}

int mpas_sw_compute_substep(mpas_ctx *ctx, double *u_provis, const double *u_old,
                            const double *u_tend, double *h_provis, const double *h_old,
                            const double *h_tend, double *tracers_provis,
                            const double *tracers_old, const double *tracers_tend,
                            double rkweight) {
  CTX_CHECK(ctx);
  NONNULL(u_provis); NONNULL(u_old); NONNULL(u_tend); NONNULL(h_provis); NONNULL(h_old);
  NONNULL(h_tend);
  const DMesh &M = ctx->dm;
  if (M.nT) { NONNULL(tracers_provis); NONNULL(tracers_old); NONNULL(tracers_tend); }
  const int nu = M.nE * M.nL;
  LAUNCH(k_axpy_edges, nu, nu, u_provis, u_old, u_tend, rkweight, 0);
  LAUNCH(k_substep_cells, M.nC, M, h_provis, h_old, h_tend, tracers_provis, tracers_old,
         tracers_tend, rkweight);
  if (ctx->params.freeze_u)  // solve.cc:307-313
    MPAS_HIP(hipMemcpyAsync(u_provis, u_old, sizeof(double) * nu, hipMemcpyDeviceToDevice,
                            ctx->stream));
  return 0;
}

int mpas_sw_accumulate_update(mpas_ctx *ctx, double *u_new, const double *u_tend,
                              double *h_new, const double *h_tend, double *tracers_new,
                              const double *tracers_tend, double rkweight) {
  CTX_CHECK(ctx);
  NONNULL(u_new); NONNULL(u_tend); NONNULL(h_new); NONNULL(h_tend);
  const DMesh &M = ctx->dm;
  if (M.nT) { NONNULL(tracers_new); NONNULL(tracers_tend); }
  const int nu = M.nE * M.nL;
  LAUNCH(k_axpy_edges, nu, nu, u_new, nullptr, u_tend, rkweight, 1);
  LAUNCH(k_accumulate_cells, M.nC, M, h_new, h_tend, tracers_new, tracers_tend, rkweight);
  return 0;
}

int mpas_sw_finalize_timestep(mpas_ctx *ctx, const double *u_old, double *u_new,
                              const double *h_new, double *tracers_new) {
  CTX_CHECK(ctx);
  NONNULL(u_old); NONNULL(u_new); NONNULL(h_new);
  const DMesh &M = ctx->dm;
  if (M.nT) {
    NONNULL(tracers_new);
    LAUNCH(k_finalize_cells, M.nC, M, h_new, tracers_new);
  }
  if (ctx->params.freeze_u)  // solve.cc:359-366
    MPAS_HIP(hipMemcpyAsync(u_new, u_old, sizeof(double) * M.nE * M.nL,
                            hipMemcpyDeviceToDevice, ctx->stream));
  return 0;
}

int mpas_sw_timeshift(mpas_ctx *ctx, double *u_old, double *h_old, double *tracers_old,
                      double *h_edge_old, double *ke_old, double *pv_edge_old,
                      const double *u_new, const double *h_new, const double *tracers_new,
                      const double *h_edge_new, const double *ke_new,
                      const double *pv_edge_new) {
  return mpas_sw_init_provis(ctx, u_new, h_new, tracers_new, h_edge_new, ke_new, pv_edge_new,
                             u_old, h_old, tracers_old, h_edge_old, ke_old, pv_edge_old);
}

int mpas_sw_compute_error(mpas_ctx *ctx, const double *h, const double *latCell,
                          const double *lonCell, double sphere_radius, double elapsed,
                          double *err) {
  CTX_CHECK(ctx);
  NONNULL(h); NONNULL(latCell); NONNULL(lonCell); NONNULL(err);
  const DMesh &M = ctx->dm;
  double *terms = nullptr;
  MPAS_HIP(hipMalloc(reinterpret_cast<void **>(&terms), sizeof(double) * 2 * M.nC));
  hipLaunchKernelGGL(k_error_terms, dim3(blocks_for(M.nC, BS)), dim3(BS), 0, ctx->stream, M, h,
                     latCell, lonCell, sphere_radius, elapsed, terms, terms + M.nC);
  int rc = after_launch("k_error_terms");
  std::vector<double> host(size_t(2) * M.nC);
  if (!rc) {
    hipError_t e = hipMemcpyAsync(host.data(), terms, sizeof(double) * 2 * M.nC,
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      set_error(std::string("mpas_sw_compute_error: ") + hipGetErrorString(e));
      rc = MPAS_EHIP;
    }
  }
  (void)hipFree(terms);
  if (rc) return rc;
  const double pi = 3.141592653589793;
  double num = 0, den = 0;
  for (int c = 0; c < M.nC; c++) {
    num += host[c];
    den += host[size_t(M.nC) + c];
  }
  num /= (4 * pi);
  den /= (4 * pi);
  *err = std::sqrt(num) / std::sqrt(den);
  return 0;
}

int mpas_sw_del2_viscosity(mpas_ctx *ctx, const double *u, const double *vorticity,
                           double *divergence, double *tend_u) {
  CTX_CHECK(ctx);
  NONNULL(u); NONNULL(vorticity); NONNULL(divergence); NONNULL(tend_u);
  const DMesh &M = ctx->dm;
  LAUNCH(k_divergence, M.nC, M, u, divergence);
  LAUNCH(k_visc, M.nE, M, divergence, vorticity, M.msd * ctx->params.visc2, tend_u);
  return 0;
}

int mpas_fct_copy_tracers(mpas_ctx *ctx, const double *tracers, double *tracersNew,
                          double *tracersPrev) {
  CTX_CHECK(ctx);
  const DMesh &M = ctx->dm;
  const size_t b = sizeof(double) * M.nC * M.nL * M.nT;
  if (!b) return 0;
  NONNULL(tracers); NONNULL(tracersNew); NONNULL(tracersPrev);
  MPAS_HIP(hipMemcpyAsync(tracersNew, tracers, b, hipMemcpyDeviceToDevice, ctx->stream));
  MPAS_HIP(hipMemcpyAsync(tracersPrev, tracers, b, hipMemcpyDeviceToDevice, ctx->stream));
  return 0;
}

int mpas_fct_reinitialize_tracer_tendencies(mpas_ctx *ctx, double *tend) {
  CTX_CHECK(ctx);
  const DMesh &M = ctx->dm;
  const int64_t n = int64_t(M.nC) * M.nL * M.nT;
  if (!n) return 0;
  NONNULL(tend);
  LAUNCH(k_fill, n, n, tend, 0.0);
  return 0;
}

int mpas_fct_calc_normal_thickness_flux(mpas_ctx *ctx, const double *layerThicknessEdge,
                                        const double *normalVelocity,
                                        double *normalThicknessFlux) {
  CTX_CHECK(ctx);
  NONNULL(layerThicknessEdge); NONNULL(normalVelocity); NONNULL(normalThicknessFlux);
  const int n = ctx->dm.nE * ctx->dm.nL;
  LAUNCH(k_ntf, n, n, layerThicknessEdge, normalVelocity, normalThicknessFlux);
  return 0;
}

int mpas_fct_calc_provisional_layer_thicknesses(mpas_ctx *ctx, const double *layerThickness,
                                                const double *normalThicknessFlux,
                                                double *hProv, double *hProvInv,
                                                double *hNewInv, double dt) {
  CTX_CHECK(ctx);
  NONNULL(layerThickness); NONNULL(normalThicknessFlux); NONNULL(hProv); NONNULL(hProvInv);
  NONNULL(hNewInv);
  const DMesh &M = ctx->dm;
  LAUNCH(k_hprov, M.nC, M, layerThickness, normalThicknessFlux, hProv, hProvInv, hNewInv, dt);
  return 0;
}

int mpas_fct_low_order(mpas_ctx *ctx, const double *normalThicknessFlux, const double *tracers,
                       double *lowOrderFlux) {
  CTX_CHECK(ctx);
  NONNULL(normalThicknessFlux); NONNULL(tracers); NONNULL(lowOrderFlux);
  const DMesh &M = ctx->dm;
  LAUNCH(k_low_order, M.nE, M, normalThicknessFlux, tracers, lowOrderFlux);
  return 0;
}

int mpas_fct_high_order(mpas_ctx *ctx, const double *normalThicknessFlux, const double *tracers,
                        double *highOrderFlux) {
  CTX_CHECK(ctx);
  if (int rc = need_fct(ctx, __func__)) return rc;
  NONNULL(normalThicknessFlux); NONNULL(tracers); NONNULL(highOrderFlux);
  const DMesh &M = ctx->dm;
  LAUNCH(k_high_order, M.nE, M, normalThicknessFlux, tracers, highOrderFlux);
  return 0;
}

int mpas_fct_correct_and_rescale(mpas_ctx *ctx, const double *hProvInv,
                                 const double *layerThickness,
                                 const double *normalThicknessFlux, const double *lowOrderFlux,
                                 const double *tracers, double *highOrderFlux, double *flxIn,
                                 double *flxOut, double *tend, double *tracerMax,
                                 double *tracerMin, double dt) {
  CTX_CHECK(ctx);
  if (int rc = need_fct(ctx, __func__)) return rc;
  NONNULL(hProvInv); NONNULL(layerThickness); NONNULL(normalThicknessFlux);
  NONNULL(lowOrderFlux); NONNULL(tracers); NONNULL(highOrderFlux); NONNULL(flxIn);
  NONNULL(flxOut); NONNULL(tend); NONNULL(tracerMax); NONNULL(tracerMin);
  const DMesh &M = ctx->dm;
  LAUNCH(k_fct_correct, M.nE, M, normalThicknessFlux, lowOrderFlux, tracers, highOrderFlux);
  LAUNCH(k_fct_cell_limit, M.nC, M, hProvInv, layerThickness, lowOrderFlux, tracers,
         highOrderFlux, flxIn, flxOut, tend, tracerMax, tracerMin, dt);
  LAUNCH(k_fct_rescale, M.nE, M, flxIn, flxOut, highOrderFlux);
  return 0;
}

int mpas_fct_accumulate_tend(mpas_ctx *ctx, const double *hProvInv, const double *layerThickness,
                             const double *highOrderFlux, const double *tracers,
                             double *tracersProv, double *tend, double dt) {
  CTX_CHECK(ctx);
  NONNULL(hProvInv); NONNULL(layerThickness); NONNULL(highOrderFlux); NONNULL(tracers);
  NONNULL(tend);
  const DMesh &M = ctx->dm;
  LAUNCH(k_fct_accumulate, M.nC, M, hProvInv, layerThickness, highOrderFlux, tracers,
         tracersProv, tend, dt);
  return 0;
}

int mpas_hmix_del2_tend(mpas_ctx *ctx, const double *layerThicknessEdge, const double *tracers,
                        double *tend, double coef) {
  CTX_CHECK(ctx);
  NONNULL(layerThicknessEdge); NONNULL(tracers); NONNULL(tend);
  const DMesh &M = ctx->dm;
  LAUNCH(k_hmix_del2, M.nC, M, layerThicknessEdge, tracers, tend, coef);
  return 0;
}

int mpas_fct_step_tracers_forward(mpas_ctx *ctx, const double *tend, const double *tracers,
                                  double *tracersNew, double *tracersPrev, double rk_weight,
                                  double rk_subweight) {
  CTX_CHECK(ctx);
  const int n = ctx->dm.nC * ctx->dm.nL * ctx->dm.nT;
  if (!n) return 0;
  NONNULL(tend); NONNULL(tracers); NONNULL(tracersNew); NONNULL(tracersPrev);
  LAUNCH(k_step_tracers, n, n, tend, tracers, tracersNew, tracersPrev, rk_weight, rk_subweight);
  return 0;
}

int mpas_buf_copy(mpas_ctx *ctx, void *dst, const void *src, size_t bytes) {
  CTX_CHECK(ctx);
  NONNULL(dst); NONNULL(src);
  if (bytes % 16 || (uintptr_t(dst) | uintptr_t(src)) % 16)
    MPAS_FAIL(MPAS_EINVAL, "mpas_buf_copy: size and addresses must be multiples of 16");
  const int64_t n = int64_t(bytes / 8), blocks = (n + BS - 1) / BS;
  if (blocks > INT32_MAX) MPAS_FAIL(MPAS_EINVAL, "mpas_buf_copy: buffer too large");
  if (n) hipLaunchKernelGGL(k_copy, dim3(unsigned(blocks)), dim3(BS), 0, ctx->stream, n,
                            static_cast<const double *>(src), static_cast<double *>(dst));
  return after_launch("k_copy");
}

int mpas_fct_update_tracers(mpas_ctx *ctx, const double *tracersNew, double *tracers) {
  CTX_CHECK(ctx);
  const size_t b = sizeof(double) * ctx->dm.nC * ctx->dm.nL * ctx->dm.nT;
  if (!b) return 0;
  NONNULL(tracersNew); NONNULL(tracers);
  MPAS_HIP(hipMemcpyAsync(tracers, tracersNew, b, hipMemcpyDeviceToDevice, ctx->stream));
  return 0;
}

}  // extern "C"
