// Stage-fused RK4 step (apps/sw/run.cc:13-70 with the FCT tracer tasks of
// apps/ocean/run.cc:19-42 as the tracer tendency).
//
// One RK stage of the reference is 6 to 14 FleCSI task launches; here it is
//   k_stage_edges  (edges)    tend_u (+ del2 viscosity), u substep/accumulate,
//                             FCT high-order minus low-order (antidiffusive) flux
//   k_stage_cells  (cells)    tend_h, h substep/accumulate, ke and divergence
//                             of the new state, FCT bounds, upwind tendency and
//                             limiter factors (or the centred tracer tendency)
//   k_stage_tracer (cells)    FCT: rescaled antidiffusive fluxes, del2 mixing,
//                             tracer substep/accumulate
//   k_diag_vertex  (vertices) vorticity and pv_vertex of the new state
//   k_diag_edge    (edges)    h_edge and pv_edge of the new state
// Each value is formed with the operations and in the order of the per-task
// path (the same device operators), so the fused step reproduces the
// reference task sequence bit for bit; the reference's copy tasks
// (init_provis, setup_timestep, timeshift) become buffer roles.
#include "ctx.h"
#include "ops.h"

namespace mpasx {

struct StageIO {
  int stage, scheme, freeze_u, use_visc, use_hmix;
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

namespace {

constexpr int BS = 256;

template <int NT>
__device__ __forceinline__ void load_tr(const double *__restrict__ p, double (&v)[NT]) {
#pragma unroll
  for (int t = 0; t < NT; t++) v[t] = p[t];
}

// ---- edges ------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(BS) void k_stage_edges(DMesh M, StageIO S) {
  const int e = gtid();
  if (e >= M.nE) return;
  const int L = M.nL;
  const int2 cc = M.e2c[e];

  // compute_tend (+ viscosity), compute_substep, accumulate_update and, at
  // the last stage, finalize_timestep for u (solve.cc:215-233, 285-289,
  // 326-330, 359-366)
  for (int k = 0; k < L; k++) {
    const size_t o = size_t(e) * L + k;
    if (S.freeze_u) {
      if (S.stage < 3) S.yu[o] = S.u0[o];
      continue;
    }
    double t = tend_u_at(M, e, k, cc, S.xu, S.h_edge, S.pv_edge, S.ke, S.xh, S.h_s);
    if (S.use_visc) t += visc_at(M, e, k, cc, M.e2v[e], S.div, S.vort, S.visc_coef);
    if (S.stage == 0) {
      const double u0 = S.u0[o];
      S.yu[o] = u0 + S.w_sub * t;
      S.nu[o] = u0 + S.w_rk * t;
    } else if (S.stage < 3) {
      S.yu[o] = S.u0[o] + S.w_sub * t;
      S.nu[o] += S.w_rk * t;
    } else {
      S.yu[o] = S.nu[o] + S.w_rk * t;
    }
  }

  if constexpr (NT > 0) {
    if (S.scheme != MPAS_TRACER_FCT) return;
    // calc_normal_thickness_flux, high-order flux and its correction by the
    // low-order flux (ocean/solve.cc:40-57, 120-167, 216-229)
    const double dv = M.dv[e];
    const int n = M.nAdv[e];
    for (int k = 0; k < L; k++) {
      const size_t o = size_t(e) * L + k;
      const double F = S.xu[o] * S.h_edge[o];
      const int mask = M.hoMask ? M.hoMask[o] : 1;
      const double wgt = F * double(mask);
      const double sgn = (F < 0.0) ? -1.0 : double(F > 0.);
      double ho[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) ho[t] = 0.0;
      for (int i = 0; i < n; i++) {
        const int ic = M.adv[size_t(i) * M.nE + e];
        if (k >= lvl_cell(M, ic)) continue;
        const double coef1 = M.advCoef[size_t(i) * M.nE + e];
        const double coef3 = M.advCoef3[size_t(i) * M.nE + e] * M.coef3rd;
        const double weightFactor = wgt * (coef1 + coef3 * sgn);
        double tv[NT];
        load_tr<NT>(S.xtr + (size_t(ic) * L + k) * NT, tv);
#pragma unroll
        for (int t = 0; t < NT; t++) ho[t] += tv[t] * weightFactor;
      }
      double *out = S.anti + o * NT;
      if (k < lvl_edge(M, e)) {
        double t1[NT], t2[NT];
        load_tr<NT>(S.xtr + (size_t(cc.x) * L + k) * NT, t1);
        load_tr<NT>(S.xtr + (size_t(cc.y) * L + k) * NT, t2);
        const double tracerWeight = double((mask + 1) & 1) * (dv * 0.5) * F;
        const double ntf1 = smax(0.0, F), ntf2 = smin(0.0, F);
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const double lo = dv * (ntf1 * t1[t] + ntf2 * t2[t]);
          double x = ho[t];
          x += tracerWeight * (t1[t] + t2[t]);
          x -= lo;
          out[t] = x;
        }
      } else {
#pragma unroll
        for (int t = 0; t < NT; t++) out[t] = ho[t];
      }
    }
  }
}

// ---- cells ------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(BS) void k_stage_cells(DMesh M, StageIO S) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int L = M.nL, C = M.nC;
  const int n = M.nEoC[c];
  const uint8_t neg = M.sgnNeg[c];
  const uint32_t perm = M.ascPerm[c];
  const double ar = M.area[c];

  for (int k = 0; k < L; k++) {
    const size_t ol = size_t(c) * L + k;
    // compute_tend, thickness part: edges in ascending id (solve.cc:191-205)
    double th = 0;
    for (int r = 0; r < n; r++) {
      const int j = (perm >> (4 * r)) & 15;
      const int e = M.c2e[j * C + c];
      const size_t oe = size_t(e) * L + k;
      const double flux = S.xu[oe] * M.dv[e] * S.h_edge[oe];
      if ((neg >> j) & 1)
        th -= flux;
      else
        th += flux;
    }
    th = th / ar;
    double hy;
    if (S.stage == 0) {
      const double h0 = S.h0[ol];
      hy = h0 + S.w_sub * th;
      S.nh[ol] = h0 + S.w_rk * th;
    } else if (S.stage < 3) {
      hy = S.h0[ol] + S.w_sub * th;
      S.nh[ol] += S.w_rk * th;
    } else {
      hy = S.nh[ol] + S.w_rk * th;
    }
    S.yh[ol] = hy;

    // ke (solve.cc:63-77) and divergence of the new state
    double ke = 0.0, dvg = 0.0;
    for (int j = 0; j < n; j++) {
      const int e = M.c2e[j * C + c];
      const double uy = S.yu[size_t(e) * L + k];
      ke += 0.25 * M.dc[e] * M.dv[e] * (uy * uy);
      if (S.use_visc) dvg += (-sign_on_cell(neg, j)) * M.dv[e] * uy;
    }
    S.ke_y[ol] = ke / ar;
    if (S.use_visc) S.div_y[ol] = dvg / ar;

    if constexpr (NT > 0) {
      double tc[NT];
      load_tr<NT>(S.xtr + ol * NT, tc);
      if (S.scheme == MPAS_TRACER_CENTRED) {
        // compute_scalar_tend (solve.cc:237-270), ascending edge id
        double te[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) te[t] = 0.0;
        for (int r = 0; r < n; r++) {
          const int j = (perm >> (4 * r)) & 15;
          const int e = M.c2e[j * C + c];
          const size_t oe = size_t(e) * L + k;
          double to[NT];
          load_tr<NT>(S.xtr + (size_t(M.cOpp[j * C + c]) * L + k) * NT, to);
          const bool first = (neg >> j) & 1;  // c == m.cells(e)[0]
          const double base = S.xu[oe] * M.dv[e] * S.h_edge[oe];
#pragma unroll
          for (int t = 0; t < NT; t++) {
            const double tracer_edge = first ? 0.5 * (tc[t] + to[t]) : 0.5 * (to[t] + tc[t]);
            const double flux = base * tracer_edge;
            if (first)
              te[t] -= flux / ar;
            else
              te[t] += flux / ar;
          }
        }
        if (S.use_hmix) {
          // tracer_hmix_del2_tend (ocean/solve.cc:328-366), file order
          const double invArea = 1.0 / ar;
          for (int j = 0; j < n; j++) {
            const int e = M.c2e[j * C + c];
            const double rho_scaling = M.msd * S.kappa * M.dv[e] / M.dc[e];
            const double lte = S.h_edge[size_t(e) * L + k];
            double to[NT];
            load_tr<NT>(S.xtr + (size_t(M.cOpp[j * C + c]) * L + k) * NT, to);
            const bool first = (neg >> j) & 1;
            const double sg = first ? -1.0 : 1.0;
#pragma unroll
            for (int t = 0; t < NT; t++) {
              const double turb = first ? to[t] - tc[t] : tc[t] - to[t];
              const double flux = lte * turb * rho_scaling;
              te[t] -= sg * flux * invArea;
            }
          }
        }
        // compute_substep / accumulate_update / finalize_timestep (tracers)
        double *yt = S.ytr + ol * NT;
        double *nt = S.ntr + ol * NT;
        if (S.stage == 0) {
          const double h0 = S.h0[ol];
          double t0[NT];
          load_tr<NT>(S.tr0 + ol * NT, t0);
#pragma unroll
          for (int t = 0; t < NT; t++) {
            yt[t] = (h0 * t0[t] + S.w_sub * te[t]) / hy;
            nt[t] = t0[t] * h0 + S.w_rk * te[t];
          }
        } else if (S.stage < 3) {
          const double h0 = S.h0[ol];
          double t0[NT];
          load_tr<NT>(S.tr0 + ol * NT, t0);
#pragma unroll
          for (int t = 0; t < NT; t++) {
            yt[t] = (h0 * t0[t] + S.w_sub * te[t]) / hy;
            nt[t] += S.w_rk * te[t];
          }
        } else {
#pragma unroll
          for (int t = 0; t < NT; t++) yt[t] = (nt[t] + S.w_rk * te[t]) / hy;
        }
      } else if (S.scheme == MPAS_TRACER_FCT) {
        // calc_provisional_layer_thicknesses (ocean/solve.cc:74-96)
        const double dtInvAreaCell = S.dt / ar;
        double hp = S.xh[ol];
        for (int j = 0; j < n; j++) {
          const int e = M.c2e[j * C + c];
          const size_t oe = size_t(e) * L + k;
          const double signedFactor = dtInvAreaCell * M.dv[e] * sign_on_cell(neg, j);
          hp += signedFactor * (S.xu[oe] * S.h_edge[oe]);
        }
        const double hProvInv = 1.0 / hp;
        // bounds over cellsOnCell (195-214)
        double tmax[NT], tmin[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) tmax[t] = tmin[t] = tc[t];
        for (int j = 0; j < n; j++) {
          double tn[NT];
          load_tr<NT>(S.xtr + (size_t(M.c2c[j * C + c]) * L + k) * NT, tn);
#pragma unroll
          for (int t = 0; t < NT; t++) {
            tmax[t] = smax(tmax[t], tn[t]);
            tmin[t] = smin(tmin[t], tn[t]);
          }
        }
        // upwind tendency and antidiffusive in/out sums, file order (232-249)
        const double invAreaCell1 = 1.0 / ar;
        double te[NT], fin[NT], fout[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) te[t] = fin[t] = fout[t] = 0.0;
        for (int j = 0; j < n; j++) {
          const int e = M.c2e[j * C + c];
          const size_t oe = size_t(e) * L + k;
          const double signedFactor = sign_on_cell(neg, j) * invAreaCell1;
          const double F = S.xu[oe] * S.h_edge[oe];
          const double ntf1 = smax(0.0, F), ntf2 = smin(0.0, F);
          const double dv = M.dv[e];
          double to[NT], an[NT];
          load_tr<NT>(S.xtr + (size_t(M.cOpp[j * C + c]) * L + k) * NT, to);
          load_tr<NT>(S.anti + oe * NT, an);
          const bool first = (neg >> j) & 1;
#pragma unroll
          for (int t = 0; t < NT; t++) {
            const double t1 = first ? tc[t] : to[t];
            const double t2 = first ? to[t] : tc[t];
            const double lo = dv * (ntf1 * t1 + ntf2 * t2);
            te[t] += signedFactor * lo;
            fin[t] += smax(0.0, signedFactor * an[t]);
            fout[t] += smin(0.0, signedFactor * an[t]);
          }
        }
        // limiter factors (253-266)
        const double eps = 1.e-10;
        const double lt = S.xh[ol];
        double *rf = S.rfac + ol * 2 * NT;
        double *tl = S.tlo + ol * NT;
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const double tracerUpwindNew = (tc[t] * lt + S.dt * te[t]) * hProvInv;
          const double tracerMaxNew = tracerUpwindNew + S.dt * fin[t] * hProvInv;
          const double tracerMinNew = tracerUpwindNew + S.dt * fout[t] * hProvInv;
          double scaleFactor =
              (tmax[t] - tracerUpwindNew) / (tracerMaxNew - tracerUpwindNew + eps);
          rf[t] = smin(1.0, smax(0.0, scaleFactor));
          scaleFactor = (tracerUpwindNew - tmin[t]) / (tracerUpwindNew - tracerMinNew + eps);
          rf[NT + t] = smin(1.0, smax(0.0, scaleFactor));
          tl[t] = te[t];
        }
      }
    }
  }
}

// ---- cells: FCT rescale + accumulate, del2 mixing, tracer update ----------
template <int NT>
__global__ __launch_bounds__(BS) void k_stage_tracer(DMesh M, StageIO S) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int L = M.nL, C = M.nC;
  const int n = M.nEoC[c];
  const uint8_t neg = M.sgnNeg[c];
  const double ar = M.area[c];
  const double invAreaCell1 = 1.0 / ar;
  for (int k = 0; k < L; k++) {
    const size_t ol = size_t(c) * L + k;
    double te[NT], rs[2 * NT];
    load_tr<NT>(S.tlo + ol * NT, te);
    load_tr<2 * NT>(S.rfac + ol * 2 * NT, rs);
    // rescale (ocean/solve.cc:270-282) and accumulate (301-316), file order
    for (int j = 0; j < n; j++) {
      const int e = M.c2e[j * C + c];
      const size_t oe = size_t(e) * L + k;
      const int opp = M.cOpp[j * C + c];
      double ro[2 * NT], an[NT];
      load_tr<2 * NT>(S.rfac + (size_t(opp) * L + k) * 2 * NT, ro);
      load_tr<NT>(S.anti + oe * NT, an);
      const bool first = (neg >> j) & 1;
      const double signedFactor = (first ? -1.0 : 1.0) * invAreaCell1;
#pragma unroll
      for (int t = 0; t < NT; t++) {
        // factors of m.cells(e)[0] and [1]: (in, out)
        const double in1 = first ? rs[t] : ro[t], out1 = first ? rs[NT + t] : ro[NT + t];
        const double in2 = first ? ro[t] : rs[t], out2 = first ? ro[NT + t] : rs[NT + t];
        const double x = an[t];
        const double hof = smax(0.0, x) * smin(out1, in2) + smin(0.0, x) * smin(in1, out2);
        te[t] += signedFactor * hof;
      }
    }
    double tc[NT];
    load_tr<NT>(S.xtr + ol * NT, tc);
    if (S.use_hmix) {
      // tracer_hmix_del2_tend (ocean/solve.cc:328-366)
      for (int j = 0; j < n; j++) {
        const int e = M.c2e[j * C + c];
        const double rho_scaling = M.msd * S.kappa * M.dv[e] / M.dc[e];
        const double lte = S.h_edge[size_t(e) * L + k];
        double to[NT];
        load_tr<NT>(S.xtr + (size_t(M.cOpp[j * C + c]) * L + k) * NT, to);
        const bool first = (neg >> j) & 1;
        const double sg = first ? -1.0 : 1.0;
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const double turb = first ? to[t] - tc[t] : tc[t] - to[t];
          const double flux = lte * turb * rho_scaling;
          te[t] -= sg * flux * invAreaCell1;
        }
      }
    }
    // compute_substep / accumulate_update / finalize_timestep (tracers)
    const double hy = S.yh[ol];
    double *yt = S.ytr + ol * NT;
    double *nt = S.ntr + ol * NT;
    if (S.stage == 0) {
      const double h0 = S.h0[ol];
      double t0[NT];
      load_tr<NT>(S.tr0 + ol * NT, t0);
#pragma unroll
      for (int t = 0; t < NT; t++) {
        yt[t] = (h0 * t0[t] + S.w_sub * te[t]) / hy;
        nt[t] = t0[t] * h0 + S.w_rk * te[t];
      }
    } else if (S.stage < 3) {
      const double h0 = S.h0[ol];
      double t0[NT];
      load_tr<NT>(S.tr0 + ol * NT, t0);
#pragma unroll
      for (int t = 0; t < NT; t++) {
        yt[t] = (h0 * t0[t] + S.w_sub * te[t]) / hy;
        nt[t] += S.w_rk * te[t];
      }
    } else {
#pragma unroll
      for (int t = 0; t < NT; t++) yt[t] = (nt[t] + S.w_rk * te[t]) / hy;
    }
  }
}

__global__ __launch_bounds__(BS) void k_diag_vertex(DMesh M, const double *__restrict__ h,
                                                    const double *__restrict__ u,
                                                    double *vorticity, double *pv_vertex) {
  const int v = gtid();
  if (v < M.nV) diag_vertex(M, v, h, u, nullptr, nullptr, vorticity, pv_vertex);
}

__global__ __launch_bounds__(BS) void k_diag_edge(DMesh M, const double *__restrict__ h,
                                                  const double *__restrict__ pv_vertex,
                                                  double *h_edge, double *pv_edge) {
  const int e = gtid();
  if (e < M.nE) diag_edge(M, e, h, pv_vertex, h_edge, pv_edge);
}

// ke and divergence of a state (the cell part of the initial diagnostics)
__global__ __launch_bounds__(BS) void k_diag_cell(DMesh M, const double *__restrict__ u,
                                                  double *ke, double *div, int use_visc) {
  const int c = gtid();
  if (c >= M.nC) return;
  const int L = M.nL, C = M.nC;
  const int n = M.nEoC[c];
  const uint8_t neg = M.sgnNeg[c];
  const double ar = M.area[c];
  for (int k = 0; k < L; k++) {
    double s = 0.0, d = 0.0;
    for (int j = 0; j < n; j++) {
      const int e = M.c2e[j * C + c];
      const double ue = u[size_t(e) * L + k];
      s += 0.25 * M.dc[e] * M.dv[e] * (ue * ue);
      if (use_visc) d += (-sign_on_cell(neg, j)) * M.dv[e] * ue;
    }
    ke[size_t(c) * L + k] = s / ar;
    if (use_visc) div[size_t(c) * L + k] = d / ar;
  }
}

// Compulsory HBM bytes of one launch: every array element the launch reads
// or writes, counted once (gathered neighbours included once).  This is
// the algorithmic traffic the roofline of each kernel is measured against.
struct Bytes {
  double edges, cells, tracer, vertex, edge_diag;
};

Bytes stage_bytes(const DMesh &M, const StageIO &S, double sum_eoe, double sum_adv,
                  double sum_eoc) {
  const double C = M.nC, E = M.nE, V = M.nV, L = M.nL, T = M.nT, d = 8.0;
  const bool fct = T > 0 && S.scheme == MPAS_TRACER_FCT;
  const bool cen = T > 0 && S.scheme == MPAS_TRACER_CENTRED;
  Bytes b{};
  // k_stage_edges
  double e = 0;
  if (S.freeze_u) {
    e += S.stage < 3 ? 2 * d * E * L : 0;
  } else {
    e += 8 * E + 2 * d * E + 1 * E + 4 * sum_eoe + d * sum_eoe;  // e2c, dc, dv, eoe, woe
    e += d * L * (3 * E + 2 * C) + d * C;                       // xu h_edge pv_edge, ke xh, h_s
    if (S.use_visc) e += 8 * E + d * L * (C + V);               // e2v, div, vort
    if (S.stage == 0) e += d * L * 2 * E;                       // nu, yu written (u0 = xu)
    else if (S.stage < 3) e += d * L * 4 * E;                   // u0, nu rw, yu
    else e += d * L * 2 * E;                                    // nu, yu
  }
  if (fct) e += 1 * E + 20 * sum_adv + d * L * T * (C + E);     // nAdv adv coefs, xtr, anti
  b.edges = e;
  // k_stage_cells
  double c = 6 * C + d * C + 4 * sum_eoc + 2 * d * E;           // nEoC sgn perm, area, c2e, dv dc
  c += d * L * (3 * E);                                         // xu h_edge yu
  c += d * L * C * (S.stage == 0 ? 3 : S.stage < 3 ? 4 : 3);    // h0/nh/yh
  c += d * L * C * (1 + (S.use_visc ? 1 : 0));                  // ke_y div_y
  if (fct) c += 4 * sum_eoc + d * L * (C + T * C + T * E + 3 * T * C);  // c2c, xh, xtr, anti, rfac tlo
  if (cen) c += 4 * sum_eoc + d * L * T * (S.stage == 0 ? 3 : S.stage < 3 ? 4 : 3) * C;
  b.cells = c;
  // k_stage_tracer
  double t = 0;
  if (fct) {
    t = 2 * C + d * C + 8 * sum_eoc;                            // nEoC sgn, area, c2e cOpp
    t += d * L * T * (C + 2 * C + E + C);                       // tlo rfac anti xtr
    t += d * L * C;                                             // yh
    t += d * L * T * C * (S.stage == 0 ? 2 : S.stage < 3 ? 4 : 2);  // tr0/ntr/ytr
    if (S.stage > 0 && S.stage < 3) t += d * L * C;             // h0
    if (S.use_hmix) t += 2 * d * E + d * L * E;                 // dv dc h_edge
  }
  b.tracer = t;
  b.vertex = 44 * V + 2 * d * V + d * E + d * L * (E + C + 2 * V);
  b.edge_diag = 16 * E + d * L * (C + V + 2 * E);
  return b;
}

template <int NT>
int launch_stage(mpas_ctx *ctx, const DMesh &M, const StageIO &S, hipStream_t st,
                 const Bytes &b) {
  if (int rc = prof_begin(ctx, "k_stage_edges", b.edges)) return rc;
  hipLaunchKernelGGL(k_stage_edges<NT>, dim3(blocks_for(M.nE, BS)), dim3(BS), 0, st, M, S);
  if (int rc = after_launch("k_stage_edges")) return rc;
  if (int rc = prof_end(ctx)) return rc;
  if (int rc = prof_begin(ctx, "k_stage_cells", b.cells)) return rc;
  hipLaunchKernelGGL(k_stage_cells<NT>, dim3(blocks_for(M.nC, BS)), dim3(BS), 0, st, M, S);
  if (int rc = after_launch("k_stage_cells")) return rc;
  if (int rc = prof_end(ctx)) return rc;
  if constexpr (NT > 0) {
    if (S.scheme == MPAS_TRACER_FCT) {
      if (int rc = prof_begin(ctx, "k_stage_tracer", b.tracer)) return rc;
      hipLaunchKernelGGL(k_stage_tracer<NT>, dim3(blocks_for(M.nC, BS)), dim3(BS), 0, st, M, S);
      if (int rc = after_launch("k_stage_tracer")) return rc;
      if (int rc = prof_end(ctx)) return rc;
    }
  }
  return 0;
}

int launch_stage_nt(mpas_ctx *ctx, const DMesh &M, const StageIO &S, hipStream_t st,
                    const Bytes &b) {
  switch (M.nT) {
    case 0: return launch_stage<0>(ctx, M, S, st, b);
    case 1: return launch_stage<1>(ctx, M, S, st, b);
    case 2: return launch_stage<2>(ctx, M, S, st, b);
    case 3: return launch_stage<3>(ctx, M, S, st, b);
    case 4: return launch_stage<4>(ctx, M, S, st, b);
    case 5: return launch_stage<5>(ctx, M, S, st, b);
    case 6: return launch_stage<6>(ctx, M, S, st, b);
    case 7: return launch_stage<7>(ctx, M, S, st, b);
    case 8: return launch_stage<8>(ctx, M, S, st, b);
  }
  MPAS_FAIL(MPAS_EINVAL, "the fused step supports at most 8 tracers");
}

}  // namespace

// Diagnostics of the state in (u, h) into ping-pong slot `slot`
// (compute_solve_diagnostics, plus the divergence the viscosity reads).
int sw_diagnostics(mpas_ctx *ctx, const double *u, const double *h, int slot) {
  const DMesh &M = ctx->dm;
  SwState &S = *ctx->sw;
  hipStream_t st = ctx->stream;
  const int use_visc = ctx->params.visc2 != 0.0;
  hipLaunchKernelGGL(k_diag_cell, dim3(blocks_for(M.nC, BS)), dim3(BS), 0, st, M, u,
                     S.ke[slot], S.div[slot], use_visc);
  if (int rc = after_launch("k_diag_cell")) return rc;
  hipLaunchKernelGGL(k_diag_vertex, dim3(blocks_for(M.nV, BS)), dim3(BS), 0, st, M, h, u,
                     S.vort[slot], S.pv_vertex);
  if (int rc = after_launch("k_diag_vertex")) return rc;
  hipLaunchKernelGGL(k_diag_edge, dim3(blocks_for(M.nE, BS)), dim3(BS), 0, st, M, h,
                     S.pv_vertex, S.h_edge[slot], S.pv_edge[slot]);
  return after_launch("k_diag_edge");
}

// nsteps of apps/sw/run.cc:13-70.
int sw_rk4_steps(mpas_ctx *ctx, double dt, int nsteps) {
  const DMesh &M = ctx->dm;
  SwState &W = *ctx->sw;
  hipStream_t st = ctx->stream;
  const mpas_params &P = ctx->params;
  const double rk_weights[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
  const double rk_substep_weights[4] = {dt / 2, dt / 2, dt, 0};
  const HostMesh &hm = ctx->host;
  const double sum_eoe = double(hm.e2e.idx.size()), sum_eoc = double(hm.c2e.idx.size());
  double sum_adv = 0;
  for (int x : hm.nAdvCellsForEdge) sum_adv += x;
  for (int step = 0; step < nsteps; step++) {
    for (int s = 0; s < 4; s++) {
      const int cur = W.cur, nxt = 1 - W.cur;
      StageIO S{};
      S.stage = s;
      S.scheme = M.nT ? P.tracer_scheme : MPAS_TRACER_NONE;
      S.freeze_u = P.freeze_u;
      S.use_visc = P.visc2 != 0.0;
      S.use_hmix = M.nT && P.kappa2 != 0.0;
      S.w_rk = rk_weights[s];
      S.w_sub = rk_substep_weights[s];
      S.dt = dt;
      S.visc_coef = M.msd * P.visc2;
      S.kappa = P.kappa2;
      S.xu = s == 0 ? W.u : W.pu[cur];
      S.xh = s == 0 ? W.h : W.ph[cur];
      S.xtr = s == 0 ? W.tr : W.ptr[cur];
      S.h_edge = W.h_edge[cur];
      S.ke = W.ke[cur];
      S.pv_edge = W.pv_edge[cur];
      S.vort = W.vort[cur];
      S.div = W.div[cur];
      S.h_s = W.h_s;
      S.u0 = W.u;
      S.h0 = W.h;
      S.tr0 = W.tr;
      S.nu = W.nu;
      S.nh = W.nh;
      S.ntr = W.ntr;
      S.yu = s < 3 ? W.pu[nxt] : W.u;
      S.yh = s < 3 ? W.ph[nxt] : W.h;
      S.ytr = s < 3 ? W.ptr[nxt] : W.tr;
      S.ke_y = W.ke[nxt];
      S.div_y = W.div[nxt];
      S.anti = W.anti;
      S.rfac = W.rfac;
      S.tlo = W.tlo;
      const Bytes b = stage_bytes(M, S, sum_eoe, sum_adv, sum_eoc);
      if (int rc = launch_stage_nt(ctx, M, S, st, b)) return rc;
      if (int rc = prof_begin(ctx, "k_diag_vertex", b.vertex)) return rc;
      hipLaunchKernelGGL(k_diag_vertex, dim3(blocks_for(M.nV, BS)), dim3(BS), 0, st, M, S.yh,
                         S.yu, W.vort[nxt], W.pv_vertex);
      if (int rc = after_launch("k_diag_vertex")) return rc;
      if (int rc = prof_end(ctx)) return rc;
      if (int rc = prof_begin(ctx, "k_diag_edge", b.edge_diag)) return rc;
      hipLaunchKernelGGL(k_diag_edge, dim3(blocks_for(M.nE, BS)), dim3(BS), 0, st, M, S.yh,
                         W.pv_vertex, W.h_edge[nxt], W.pv_edge[nxt]);
      if (int rc = after_launch("k_diag_edge")) return rc;
      if (int rc = prof_end(ctx)) return rc;
      W.cur = nxt;
    }
    W.elapsed += dt;
  }
  return 0;
}

}  // namespace mpasx
