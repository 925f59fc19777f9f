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
#include "stage.h"

namespace mpasx {

namespace {

constexpr int BS = 256;

template <int NT>
__device__ __forceinline__ void load_tr(const double *__restrict__ p, uint32_t i0,
                                        double (&v)[NT]) {
#pragma unroll
  for (int t = 0; t < NT; t++) v[t] = ld(p, i0 + t);
}

// TPL consecutive tracers of one entity: one 16-byte load / store per pair
// (the offset i is even whenever TPL == 2: NT is even then).
template <int TPL>
__device__ __forceinline__ void ld_group(const double *__restrict__ p, uint32_t i,
                                         double (&v)[TPL]) {
  if constexpr (TPL == 2) {
    const double2 x = ld(reinterpret_cast<const double2 *>(p), i / 2);
    v[0] = x.x;
    v[1] = x.y;
  } else {
#pragma unroll
    for (int g = 0; g < TPL; g++) v[g] = ld(p, i + g);
  }
}
template <int TPL>
__device__ __forceinline__ void st_group(double *p, uint32_t i, const double (&v)[TPL]) {
  if constexpr (TPL == 2) {
    st(reinterpret_cast<double2 *>(p), i / 2, make_double2(v[0], v[1]));
  } else {
#pragma unroll
    for (int g = 0; g < TPL; g++) st(p, i + g, v[g]);
  }
}

// Tracers handled by one lane: pairs when the count is even.
template <int NT>
constexpr int tracers_per_lane() {
  return NT % 2 == 0 ? 2 : 1;
}

// Adds the live slots' values in ascending edge id, the order of the
// reference's scatter loops: at step r the slot of rank r is selected with
// compile-time indices only (a run-time index into a register array would
// send the array to scratch memory).  acc -= v where the cell is the edge's
// first cell (negative edge sign), acc += v otherwise.
template <int W>
__device__ __forceinline__ double sum_ascending(double acc, const double (&v)[W],
                                                uint32_t rank, uint8_t neg, int n) {
#pragma unroll
  for (int r = 0; r < W; r++) {
    double f = 0.0;
    bool ng = false;
#pragma unroll
    for (int j = 0; j < W; j++) {
      const bool hit = ((rank >> (4 * j)) & 15) == uint32_t(r);
      f = hit ? v[j] : f;
      ng = hit ? bool((neg >> j) & 1) : ng;
    }
    const double x = ng ? acc - f : acc + f;
    acc = (r < n) ? x : acc;
  }
  return acc;
}

// The stage kernels read every slot of a ragged list up front, with widths
// fixed at compile time (WE edgesOnEdge, WA advection cells, W edges of a
// cell): all index loads issue together, then all gathers, so a lane waits
// for two memory round trips instead of two per slot.  Slots past a row's
// count read the ELL padding (valid ids) and are masked out of every sum,
// which still runs over the live slots in the reference's order.
//
// Tracer work runs one lane per (entity, tracer): lane i handles tracer
// i % NT of entity i / NT, so the lanes of an entity read and write its
// tracer record as one contiguous run, and each lane keeps one tracer's
// slot arrays in registers.

// ---- edges: momentum ---------------------------------------------------------
// compute_tend (+ viscosity), compute_substep, accumulate_update and, at the
// last stage, finalize_timestep for u (solve.cc:215-233, 285-289, 326-330,
// 359-366)
template <int WE>
__global__ __launch_bounds__(BS) void k_stage_edges(DMesh M, StageIO S) {
  const int e = xcd_item();
  if (e >= M.nE) return;
  const uint32_t L = M.nL, E = M.nE;
  if (S.freeze_u) {
    if (S.stage < 3)
      for (uint32_t k = 0; k < L; k++) st(S.yu, e * L + k, ld(S.u0, e * L + k));
    return;
  }
  const int2 cc = ld(M.e2c, e);
  const int n = ld(M.nEoE, e);
  uint32_t id[WE];
  double w[WE];
#pragma unroll
  for (int j = 0; j < WE; j++) {
    id[j] = ld(M.eoe, j * E + e);
    w[j] = ld(M.woe, j * E + e);
  }
  const double dc = ld(M.dc, e);
  for (uint32_t k = 0; k < L; k++) {
    const uint32_t o = e * L + k;
    double uu[WE], pv[WE], he[WE];
#pragma unroll
    for (int j = 0; j < WE; j++) {
      const uint32_t oj = id[j] * L + k;
      uu[j] = ld(S.xu, oj);
      pv[j] = ld(S.pv_edge, oj);
      he[j] = ld(S.h_edge, oj);
    }
    const double pve = ld(S.pv_edge, o);
    double q = 0;
#pragma unroll
    for (int j = 0; j < WE; j++) {
      const double workpv = 0.5 * (pve + pv[j]);
      const double qn = q + w[j] * uu[j] * workpv * he[j];
      q = (j < n) ? qn : q;
    }
    const uint32_t o1 = uint32_t(cc.x) * L + k, o2 = uint32_t(cc.y) * L + k;
    double t = q - (ld(S.ke, o2) - ld(S.ke, o1) +
                    M.gravity * (ld(S.xh, o2) + ld(S.h_s, cc.y) - ld(S.xh, o1) -
                                 ld(S.h_s, cc.x))) /
                       dc;
    if (S.use_visc) t += visc_at(M, e, k, cc, ld(M.e2v, e), S.div, S.vort, S.visc_coef);
    if (S.stage == 0) {
      const double u0 = ld(S.u0, o);
      st(S.yu, o, u0 + S.w_sub * t);
      st(S.nu, o, u0 + S.w_rk * t);
    } else if (S.stage < 3) {
      st(S.yu, o, ld(S.u0, o) + S.w_sub * t);
      st(S.nu, o, ld(S.nu, o) + S.w_rk * t);
    } else {
      st(S.yu, o, ld(S.nu, o) + S.w_rk * t);
    }
  }
}

// ---- edges x tracers: FCT high-order minus low-order flux -------------------
// calc_normal_thickness_flux, the high-order flux and its correction by the
// low-order flux (ocean/solve.cc:40-57, 120-167, 216-229); TPL tracers per
// lane.
template <int NT, int WA>
__global__ __launch_bounds__(BS) void k_stage_ho(DMesh M, StageIO S) {
  constexpr int TPL = tracers_per_lane<NT>(), NPL = NT / TPL;
  const int item = xcd_item();
  const int e = item / NPL, t0 = (item % NPL) * TPL;
  if (e >= M.nE) return;
  const uint32_t L = M.nL, E = M.nE;
  const int2 cc = ld(M.e2c, e);
  const double dv = ld(M.dv, e);
  const int na = ld(M.nAdv, e);
  uint32_t ic[WA];
  double c1[WA], c3[WA];
#pragma unroll
  for (int i = 0; i < WA; i++) {
    ic[i] = ld(M.adv, i * E + e);
    c1[i] = ld(M.advCoef, i * E + e);
    c3[i] = ld(M.advCoef3, i * E + e) * M.coef3rd;
  }
  for (uint32_t k = 0; k < L; k++) {
    const uint32_t o = e * L + k;
    const double F = ld(S.xu, o) * ld(S.h_edge, o);
    const int mask = M.hoMask ? ld(M.hoMask, o) : 1;
    const double wgt = F * double(mask);
    const double sgn = (F < 0.0) ? -1.0 : double(F > 0.);
    double tv[WA][TPL];
#pragma unroll
    for (int i = 0; i < WA; i++) ld_group<TPL>(S.xtr, (ic[i] * L + k) * NT + t0, tv[i]);
    double ho[TPL];
#pragma unroll
    for (int g = 0; g < TPL; g++) ho[g] = 0.0;
#pragma unroll
    for (int i = 0; i < WA; i++) {
      const double weightFactor = wgt * (c1[i] + c3[i] * sgn);
      const bool live = i < na && int(k) < lvl_cell(M, ic[i]);
#pragma unroll
      for (int g = 0; g < TPL; g++) {
        const double x = ho[g] + tv[i][g] * weightFactor;
        ho[g] = live ? x : ho[g];
      }
    }
    if (int(k) < lvl_edge(M, e)) {
      double t1[TPL], t2[TPL];
      ld_group<TPL>(S.xtr, (uint32_t(cc.x) * L + k) * NT + t0, t1);
      ld_group<TPL>(S.xtr, (uint32_t(cc.y) * L + k) * NT + t0, t2);
      const double tracerWeight = double((mask + 1) & 1) * (dv * 0.5) * F;
      const double ntf1 = smax(0.0, F), ntf2 = smin(0.0, F);
#pragma unroll
      for (int g = 0; g < TPL; g++) {
        const double lo = dv * (ntf1 * t1[g] + ntf2 * t2[g]);
        ho[g] += tracerWeight * (t1[g] + t2[g]);
        ho[g] -= lo;
      }
    }
    st_group<TPL>(S.anti, o * NT + t0, ho);
  }
}

// compute_substep / accumulate_update / finalize_timestep for one tracer of
// one cell (solve.cc:297-305, 332-340, 350-357)
__device__ __forceinline__ void tracer_update(const StageIO &S, uint32_t ol, uint32_t ot,
                                              double te, double hy) {
  if (S.stage == 0) {
    const double h0 = ld(S.h0, ol), t0 = ld(S.tr0, ot);
    st(S.ytr, ot, (h0 * t0 + S.w_sub * te) / hy);
    st(S.ntr, ot, t0 * h0 + S.w_rk * te);
  } else if (S.stage < 3) {
    st(S.ytr, ot, (ld(S.h0, ol) * ld(S.tr0, ot) + S.w_sub * te) / hy);
    st(S.ntr, ot, ld(S.ntr, ot) + S.w_rk * te);
  } else {
    st(S.ytr, ot, (ld(S.ntr, ot) + S.w_rk * te) / hy);
  }
}

// ---- cells (x tracers) -------------------------------------------------------
// Thickness tendency and update, ke and divergence of the new state (the
// cell's first lane stores them), and per tracer either the centred
// tendency with its update or the FCT bounds, upwind tendency and limiter
// factors.  NPL lanes per cell, TPL tracers each (one lane without tracers).
template <int NT, int W>
__global__ __launch_bounds__(BS) void k_stage_cells(DMesh M, StageIO S) {
  constexpr int TPL = NT > 0 ? tracers_per_lane<NT>() : 1;
  constexpr int NPL = NT > 0 ? NT / TPL : 1;
  const int item = xcd_item();
  const int c = item / NPL, t0 = (item % NPL) * TPL;
  if (c >= M.nC) return;
  const uint32_t L = M.nL, C = M.nC;
  const int n = ld(M.nEoC, c);
  const uint8_t neg = ld(M.sgnNeg, c);
  const uint32_t rank = ld(M.ascRank, c);
  const double ar = ld(M.area, c);
  const bool lead = t0 == 0;
  uint32_t e[W];
  double dv[W];
#pragma unroll
  for (int j = 0; j < W; j++) e[j] = ld(M.c2e, j * C + c);
#pragma unroll
  for (int j = 0; j < W; j++) dv[j] = ld(M.dv, e[j]);

  for (uint32_t k = 0; k < L; k++) {
    const uint32_t ol = c * L + k;
    if (lead) {
      // ke (solve.cc:63-77) and divergence of the new state
      double yu[W], dc[W];
#pragma unroll
      for (int j = 0; j < W; j++) {
        yu[j] = ld(S.yu, e[j] * L + k);
        dc[j] = ld(M.dc, e[j]);
      }
      double ke = 0.0, dvg = 0.0;
#pragma unroll
      for (int j = 0; j < W; j++) {
        const double kx = ke + 0.25 * dc[j] * dv[j] * (yu[j] * yu[j]);
        const double dx = dvg + (-sign_on_cell(neg, j)) * dv[j] * yu[j];
        ke = (j < n) ? kx : ke;
        dvg = (j < n) ? dx : dvg;
      }
      st(S.ke_y, ol, ke / ar);
      if (S.use_visc) st(S.div_y, ol, dvg / ar);
    }
    double xu[W], he[W];
#pragma unroll
    for (int j = 0; j < W; j++) {
      xu[j] = ld(S.xu, e[j] * L + k);
      he[j] = ld(S.h_edge, e[j] * L + k);
    }
    // compute_tend, thickness part: edges in ascending id (solve.cc:191-205)
    double th;
    {
      double fl[W];
#pragma unroll
      for (int j = 0; j < W; j++) fl[j] = xu[j] * dv[j] * he[j];
      th = sum_ascending<W>(0.0, fl, rank, neg, n);
    }
    th = th / ar;
    double hy;
    if (S.stage == 0) {
      const double h0 = ld(S.h0, ol);
      hy = h0 + S.w_sub * th;
      if (lead) st(S.nh, ol, h0 + S.w_rk * th);
    } else if (S.stage < 3) {
      hy = ld(S.h0, ol) + S.w_sub * th;
      if (lead) st(S.nh, ol, ld(S.nh, ol) + S.w_rk * th);
    } else {
      hy = ld(S.nh, ol) + S.w_rk * th;
    }
    if (lead) st(S.yh, ol, hy);

    if constexpr (NT > 0) {
      const double invArea = 1.0 / ar;
      const uint32_t ot = ol * NT + t0;
      double tc[TPL];
      ld_group<TPL>(S.xtr, ot, tc);
      double tn[W][TPL];
      {
        uint32_t op[W];
#pragma unroll
        for (int j = 0; j < W; j++) op[j] = ld(M.cOpp, j * C + c);
#pragma unroll
        for (int j = 0; j < W; j++) ld_group<TPL>(S.xtr, (op[j] * L + k) * NT + t0, tn[j]);
      }
      if (S.scheme == MPAS_TRACER_CENTRED) {
        // compute_scalar_tend (solve.cc:237-270), ascending edge id, then
        // tracer_hmix_del2_tend (ocean/solve.cc:328-366), file order
#pragma unroll
        for (int g = 0; g < TPL; g++) {
          double term[W];
#pragma unroll
          for (int j = 0; j < W; j++) {
            const bool first = (neg >> j) & 1;
            const double tracer_edge =
                first ? 0.5 * (tc[g] + tn[j][g]) : 0.5 * (tn[j][g] + tc[g]);
            term[j] = (xu[j] * dv[j] * he[j] * tracer_edge) / ar;
          }
          double te = sum_ascending<W>(0.0, term, rank, neg, n);
          if (S.use_hmix) {
#pragma unroll
            for (int j = 0; j < W; j++) {
              const bool first = (neg >> j) & 1;
              const double rho_scaling = M.msd * S.kappa * dv[j] / ld(M.dc, e[j]);
              const double turb = first ? tn[j][g] - tc[g] : tc[g] - tn[j][g];
              const double flux = he[j] * turb * rho_scaling;
              const double x = te - (first ? -1.0 : 1.0) * flux * invArea;
              te = (j < n) ? x : te;
            }
          }
          tracer_update(S, ol, ot + g, te, hy);
        }
      } else if (S.scheme == MPAS_TRACER_FCT) {
        // calc_provisional_layer_thicknesses (ocean/solve.cc:74-96)
        const double dtInvAreaCell = S.dt / ar;
        const double lt = ld(S.xh, ol);
        double hp = lt;
#pragma unroll
        for (int j = 0; j < W; j++) {
          const double signedFactor = dtInvAreaCell * dv[j] * sign_on_cell(neg, j);
          const double x = hp + signedFactor * (xu[j] * he[j]);
          hp = (j < n) ? x : hp;
        }
        const double hProvInv = 1.0 / hp;
        double an[W][TPL];
#pragma unroll
        for (int j = 0; j < W; j++) ld_group<TPL>(S.anti, (e[j] * L + k) * NT + t0, an[j]);
        double rf[2 * TPL], tl[TPL];
#pragma unroll
        for (int g = 0; g < TPL; g++) {
          // bounds over cellsOnCell (195-214); upwind tendency and
          // antidiffusive in/out sums, file order (232-249)
          double tmax = tc[g], tmin = tc[g], te = 0.0, fin = 0.0, fout = 0.0;
#pragma unroll
          for (int j = 0; j < W; j++) {
            const bool live = j < n;
            tmax = live ? smax(tmax, tn[j][g]) : tmax;
            tmin = live ? smin(tmin, tn[j][g]) : tmin;
            const double signedFactor = sign_on_cell(neg, j) * invArea;
            const double F = xu[j] * he[j];
            const double ntf1 = smax(0.0, F), ntf2 = smin(0.0, F);
            const bool first = (neg >> j) & 1;
            const double t1 = first ? tc[g] : tn[j][g];
            const double t2 = first ? tn[j][g] : tc[g];
            const double lo = dv[j] * (ntf1 * t1 + ntf2 * t2);
            const double xe = te + signedFactor * lo;
            const double xi = fin + smax(0.0, signedFactor * an[j][g]);
            const double xo = fout + smin(0.0, signedFactor * an[j][g]);
            te = live ? xe : te;
            fin = live ? xi : fin;
            fout = live ? xo : fout;
          }
          // limiter factors (253-266), stored (in, out) per tracer
          const double eps = 1.e-10;
          const double tracerUpwindNew = (tc[g] * lt + S.dt * te) * hProvInv;
          const double tracerMaxNew = tracerUpwindNew + S.dt * fin * hProvInv;
          const double tracerMinNew = tracerUpwindNew + S.dt * fout * hProvInv;
          double scaleFactor =
              (tmax - tracerUpwindNew) / (tracerMaxNew - tracerUpwindNew + eps);
          rf[2 * g] = smin(1.0, smax(0.0, scaleFactor));
          scaleFactor = (tracerUpwindNew - tmin) / (tracerUpwindNew - tracerMinNew + eps);
          rf[2 * g + 1] = smin(1.0, smax(0.0, scaleFactor));
          tl[g] = te;
        }
        if (!S.keep_rfac) {
#pragma unroll
          for (int g = 0; g < TPL; g++) {
            const double pair[2] = {rf[2 * g], rf[2 * g + 1]};
            st_group<2>(S.rfac, (ot + g) * 2, pair);
          }
        }
        st_group<TPL>(S.tlo, ot, tl);
      }
    }
  }
}

// ---- cells x tracers: FCT rescale + accumulate, del2 mixing, update --------
template <int NT, int W>
__global__ __launch_bounds__(BS) void k_stage_tracer(DMesh M, StageIO S) {
  constexpr int TPL = tracers_per_lane<NT>(), NPL = NT / TPL;
  const int item = xcd_item();
  const int c = item / NPL, t0 = (item % NPL) * TPL;
  if (c >= M.nC) return;
  const uint32_t L = M.nL, C = M.nC;
  const int n = ld(M.nEoC, c);
  const uint8_t neg = ld(M.sgnNeg, c);
  const double invAreaCell1 = 1.0 / ld(M.area, c);
  uint32_t e[W], op[W];
#pragma unroll
  for (int j = 0; j < W; j++) {
    e[j] = ld(M.c2e, j * C + c);
    op[j] = ld(M.cOpp, j * C + c);
  }
  for (uint32_t k = 0; k < L; k++) {
    const uint32_t ol = c * L + k, ot = ol * NT + t0;
    double te[TPL], rs[2 * TPL];
    ld_group<TPL>(S.tlo, ot, te);
#pragma unroll
    for (int g = 0; g < TPL; g++) {
      double pair[2];
      ld_group<2>(S.rfac, (ot + g) * 2, pair);
      rs[2 * g] = pair[0];
      rs[2 * g + 1] = pair[1];
    }
    double an[W][TPL], ro[W][2 * TPL];
#pragma unroll
    for (int j = 0; j < W; j++) {
      ld_group<TPL>(S.anti, (e[j] * L + k) * NT + t0, an[j]);
      const uint32_t oo = (op[j] * L + k) * NT + t0;
#pragma unroll
      for (int g = 0; g < TPL; g++) {
        double pair[2];
        ld_group<2>(S.rfac, (oo + g) * 2, pair);
        ro[j][2 * g] = pair[0];
        ro[j][2 * g + 1] = pair[1];
      }
    }
    // rescale (ocean/solve.cc:270-282) and accumulate (301-316), file order
#pragma unroll
    for (int j = 0; j < W; j++) {
      const bool first = (neg >> j) & 1;  // c == m.cells(e)[0]
      const double signedFactor = (first ? -1.0 : 1.0) * invAreaCell1;
#pragma unroll
      for (int g = 0; g < TPL; g++) {
        const double rin = rs[2 * g], rout = rs[2 * g + 1];
        const double oin = ro[j][2 * g], oout = ro[j][2 * g + 1];
        const double in1 = first ? rin : oin, out1 = first ? rout : oout;
        const double in2 = first ? oin : rin, out2 = first ? oout : rout;
        const double x = an[j][g];
        const double hof = smax(0.0, x) * smin(out1, in2) + smin(0.0, x) * smin(in1, out2);
        const double y = te[g] + signedFactor * hof;
        te[g] = (j < n) ? y : te[g];
      }
    }
    if (S.use_hmix) {
      // tracer_hmix_del2_tend (ocean/solve.cc:328-366)
      double tc[TPL], tn[W][TPL], he[W], rho[W];
      ld_group<TPL>(S.xtr, ot, tc);
#pragma unroll
      for (int j = 0; j < W; j++) {
        ld_group<TPL>(S.xtr, (op[j] * L + k) * NT + t0, tn[j]);
        he[j] = ld(S.h_edge, e[j] * L + k);
        rho[j] = M.msd * S.kappa * ld(M.dv, e[j]) / ld(M.dc, e[j]);
      }
#pragma unroll
      for (int j = 0; j < W; j++) {
        const bool first = (neg >> j) & 1;
#pragma unroll
        for (int g = 0; g < TPL; g++) {
          const double turb = first ? tn[j][g] - tc[g] : tc[g] - tn[j][g];
          const double flux = he[j] * turb * rho[j];
          const double y = te[g] - (first ? -1.0 : 1.0) * flux * invAreaCell1;
          te[g] = (j < n) ? y : te[g];
        }
      }
    }
    const double hy = ld(S.yh, ol);
#pragma unroll
    for (int g = 0; g < TPL; g++) tracer_update(S, ol, ot + g, te[g], hy);
  }
}

__global__ __launch_bounds__(BS) void k_diag_vertex(DMesh M, const double *__restrict__ h,
                                                    const double *__restrict__ u,
                                                    double *vorticity, double *pv_vertex) {
  const int v = xcd_item();
  if (v < M.nV) diag_vertex(M, v, h, u, nullptr, nullptr, vorticity, pv_vertex);
}

__global__ __launch_bounds__(BS) void k_diag_edge(DMesh M, const double *__restrict__ h,
                                                  const double *__restrict__ pv_vertex,
                                                  double *h_edge, double *pv_edge) {
  const int e = xcd_item();
  if (e < M.nE) diag_edge(M, e, h, pv_vertex, h_edge, pv_edge);
}

// ke and divergence of a state (the cell part of the initial diagnostics)
__global__ __launch_bounds__(BS) void k_diag_cell(DMesh M, const double *__restrict__ u,
                                                  double *ke, double *div, int use_visc) {
  const int c = xcd_item();
  if (c >= M.nC) return;
  const uint32_t L = M.nL, C = M.nC;
  const int n = ld(M.nEoC, c);
  const uint8_t neg = ld(M.sgnNeg, c);
  const double ar = ld(M.area, c);
  for (uint32_t k = 0; k < L; k++) {
    double s = 0.0, d = 0.0;
    for (int j = 0; j < n; j++) {
      const uint32_t e = ld(M.c2e, j * C + c);
      const double ue = ld(u, e * L + k);
      s += 0.25 * ld(M.dc, e) * ld(M.dv, e) * (ue * ue);
      if (use_visc) d += (-sign_on_cell(neg, j)) * ld(M.dv, e) * ue;
    }
    st(ke, c * L + k, s / ar);
    if (use_visc) st(div, c * L + k, d / ar);
  }
}

// Compulsory HBM bytes of one launch: every array element the launch reads
// or writes, counted once (gathered neighbours included once).  This is
// the algorithmic traffic the roofline of each kernel is measured against.
struct Bytes {
  double edges, ho, cells, tracer, vertex, edge_diag;
};

Bytes stage_bytes(const DMesh &M, const StageIO &S, double sum_eoe, double sum_adv,
                  double sum_eoc) {
  const double C = M.nC, E = M.nE, V = M.nV, L = M.nL, T = M.nT, d = 8.0;
  const bool fct = T > 0 && S.scheme == MPAS_TRACER_FCT;
  const bool cen = T > 0 && S.scheme == MPAS_TRACER_CENTRED;
  // state touches of an RK4 stage: stage 0 reads prev (= provis) and writes
  // provis and next; stages 1-2 read prev, read+write next, write provis;
  // stage 3 reads next and writes the final state
  const double touches = S.stage == 0 ? 2 : S.stage < 3 ? 4 : 2;
  Bytes b{};
  // k_stage_edges: momentum
  if (S.freeze_u) {
    b.edges = S.stage < 3 ? 2 * d * L * E : 0;
  } else {
    b.edges = 8 * E + d * E + 1 * E + 4 * sum_eoe + d * sum_eoe;   // e2c, dc, eoe, woe
    b.edges += d * L * (3 * E + 2 * C) + d * C;                    // xu h_edge pv_edge, ke xh, h_s
    if (S.use_visc) b.edges += 8 * E + d * E + d * L * (C + V);    // e2v, dv, div, vort
    b.edges += touches * d * L * E;                                // u0 / nu / yu
  }
  // k_stage_ho: antidiffusive flux per edge and tracer
  if (fct) b.ho = 8 * E + d * E + E + 20 * sum_adv + 2 * d * L * E + d * L * T * (C + E);
  // k_stage_cells
  b.cells = 14 * C + 4 * sum_eoc + 2 * d * E;                      // nEoC sgn rank area, c2e, dv dc
  b.cells += d * L * 3 * E;                                        // yu xu h_edge
  b.cells += d * L * C * (touches + 1 + (S.use_visc ? 1 : 0));     // h0/nh/yh, ke_y, div_y
  if (fct) b.cells += 4 * sum_eoc + d * L * (C + T * C + T * E + 3 * T * C);  // cOpp, xh, xtr, anti, rfac tlo
  if (cen) b.cells += 4 * sum_eoc + d * L * T * C * (touches + 1);  // cOpp, xtr + tr0/ntr/ytr
  // k_stage_tracer
  if (fct) {
    b.tracer = 10 * C + 8 * sum_eoc;                               // nEoC sgn area, c2e cOpp
    b.tracer += d * L * T * (C + 2 * C + E) + d * L * C;           // tlo rfac anti, yh
    b.tracer += d * L * T * C * touches + (S.stage > 0 && S.stage < 3 ? d * L * C : 0);
    if (S.use_hmix) b.tracer += 2 * d * E + d * L * E + d * L * T * C;  // dv dc h_edge xtr
  }
  b.vertex = 44 * V + 2 * d * V + d * E + d * L * (E + C + 2 * V);
  b.edge_diag = 16 * E + d * L * (C + V + 2 * E);
  return b;
}

// One RK stage, in order on the context's stream: momentum, high-order
// flux (FCT), cells, tracer update (FCT), then the diagnostics of the
// stage's output.  (Running the independent pairs on two streams at once was
// measured slower on the level-8 workload: the pairs contend for L2 and HBM.)
template <int NT, int W, int WE>
int launch_stage_t(mpas_ctx *ctx, const DMesh &M, const StageIO &S, const Bytes &b,
                 double *vort_y, double *pv_vertex, double *h_edge_y, double *pv_edge_y) {
  constexpr int NPL = NT > 0 ? NT / tracers_per_lane<NT>() : 1;  // lanes per entity
  const bool fct = NT > 0 && S.scheme == MPAS_TRACER_FCT;
  hipStream_t st = ctx->stream;
#define MPAS_STEP(expr) \
  do {                  \
    if (int rc = (expr)) return rc; \
  } while (0)
  MPAS_STEP(prof_begin(ctx, "k_stage_edges", b.edges));
  hipLaunchKernelGGL((k_stage_edges<WE>), dim3(xcd_grid(M.nE, BS)), dim3(BS), 0, st, M, S);
  MPAS_STEP(after_launch("k_stage_edges"));
  MPAS_STEP(prof_end(ctx));
  if constexpr (NT > 0) {
    if (fct) {
      MPAS_STEP(prof_begin(ctx, "k_stage_ho", b.ho));
      hipLaunchKernelGGL((k_stage_ho<NT, WE>), dim3(xcd_grid(int64_t(M.nE) * NPL, BS)), dim3(BS),
                         0, st, M, S);
      MPAS_STEP(after_launch("k_stage_ho"));
      MPAS_STEP(prof_end(ctx));
    }
  }
  MPAS_STEP(prof_begin(ctx, "k_stage_cells", b.cells));
  hipLaunchKernelGGL((k_stage_cells<NT, W>), dim3(xcd_grid(int64_t(M.nC) * NPL, BS)), dim3(BS),
                     0, st, M, S);
  MPAS_STEP(after_launch("k_stage_cells"));
  MPAS_STEP(prof_end(ctx));
  if constexpr (NT > 0) {
    if (fct) {
      MPAS_STEP(prof_begin(ctx, "k_stage_tracer", b.tracer));
      hipLaunchKernelGGL((k_stage_tracer<NT, W>), dim3(xcd_grid(int64_t(M.nC) * NPL, BS)),
                         dim3(BS), 0, st, M, S);
      MPAS_STEP(after_launch("k_stage_tracer"));
      MPAS_STEP(prof_end(ctx));
    }
  }
  MPAS_STEP(prof_begin(ctx, "k_diag_vertex", b.vertex));
  hipLaunchKernelGGL(k_diag_vertex, dim3(xcd_grid(M.nV, BS)), dim3(BS), 0, st, M, S.yh, S.yu,
                     vort_y, pv_vertex);
  MPAS_STEP(after_launch("k_diag_vertex"));
  MPAS_STEP(prof_end(ctx));
  MPAS_STEP(prof_begin(ctx, "k_diag_edge", b.edge_diag));
  hipLaunchKernelGGL(k_diag_edge, dim3(xcd_grid(M.nE, BS)), dim3(BS), 0, st, M, S.yh, pv_vertex,
                     h_edge_y, pv_edge_y);
  MPAS_STEP(after_launch("k_diag_edge"));
  MPAS_STEP(prof_end(ctx));
#undef MPAS_STEP
  return 0;
}

#define MPAS_STAGE_ARGS ctx, M, S, b, vort_y, pv_vertex, h_edge_y, pv_edge_y
template <int NT>
int launch_stage_w(mpas_ctx *ctx, const DMesh &M, const StageIO &S, const Bytes &b,
                   double *vort_y, double *pv_vertex, double *h_edge_y, double *pv_edge_y) {
  if (M.narrow) return launch_stage_t<NT, 6, 10>(MPAS_STAGE_ARGS);
  return launch_stage_t<NT, 8, 16>(MPAS_STAGE_ARGS);
}

int launch_stage_nt(mpas_ctx *ctx, const DMesh &M, const StageIO &S, const Bytes &b,
                    double *vort_y, double *pv_vertex, double *h_edge_y, double *pv_edge_y) {
  switch (M.nT) {
    case 0: return launch_stage_w<0>(MPAS_STAGE_ARGS);
    case 1: return launch_stage_w<1>(MPAS_STAGE_ARGS);
    case 2: return launch_stage_w<2>(MPAS_STAGE_ARGS);
    case 3: return launch_stage_w<3>(MPAS_STAGE_ARGS);
    case 4: return launch_stage_w<4>(MPAS_STAGE_ARGS);
    case 5: return launch_stage_w<5>(MPAS_STAGE_ARGS);
    case 6: return launch_stage_w<6>(MPAS_STAGE_ARGS);
    case 7: return launch_stage_w<7>(MPAS_STAGE_ARGS);
    case 8: return launch_stage_w<8>(MPAS_STAGE_ARGS);
  }
  MPAS_FAIL(MPAS_EINVAL, "the fused step supports at most 8 tracers");
}

}  // namespace

StageIO stage_params(const mpas_ctx *ctx, int s, double dt) {
  const DMesh &M = ctx->dm;
  const mpas_params &P = ctx->params;
  StageIO S{};
  S.stage = s;
  S.scheme = M.nT ? P.tracer_scheme : MPAS_TRACER_NONE;
  S.freeze_u = P.freeze_u;
  S.use_visc = P.visc2 != 0.0;
  S.use_hmix = M.nT && P.kappa2 != 0.0;
  S.w_rk = rk_weight(s, dt);
  S.w_sub = rk_substep_weight(s, dt);
  S.dt = dt;
  S.visc_coef = M.msd * P.visc2;
  S.kappa = P.kappa2;
  return S;
}

int launch_stage(mpas_ctx *ctx, const StageIO &S, double *vort_y, double *pv_vertex,
                 double *h_edge_y, double *pv_edge_y) {
  const DMesh &M = ctx->dm;
  const HostMesh &hm = ctx->host;
  const double sum_eoe = double(hm.e2e.idx.size()), sum_eoc = double(hm.c2e.idx.size());
  double sum_adv = 0;
  if (S.scheme == MPAS_TRACER_FCT)
    for (int x : hm.nAdvCellsForEdge) sum_adv += x;
  const Bytes b = stage_bytes(M, S, sum_eoe, sum_adv, sum_eoc);
  return launch_stage_nt(ctx, M, S, b, vort_y, pv_vertex, h_edge_y, pv_edge_y);
}

int diagnostics_into(mpas_ctx *ctx, const double *u, const double *h, double *ke, double *div,
                     double *vort, double *pv_vertex, double *h_edge, double *pv_edge) {
  const DMesh &M = ctx->dm;
  hipStream_t st = ctx->stream;
  const int use_visc = ctx->params.visc2 != 0.0;
  hipLaunchKernelGGL(k_diag_cell, dim3(xcd_grid(M.nC, BS)), dim3(BS), 0, st, M, u, ke, div,
                     use_visc);
  if (int rc = after_launch("k_diag_cell")) return rc;
  hipLaunchKernelGGL(k_diag_vertex, dim3(xcd_grid(M.nV, BS)), dim3(BS), 0, st, M, h, u, vort,
                     pv_vertex);
  if (int rc = after_launch("k_diag_vertex")) return rc;
  hipLaunchKernelGGL(k_diag_edge, dim3(xcd_grid(M.nE, BS)), dim3(BS), 0, st, M, h, pv_vertex,
                     h_edge, pv_edge);
  return after_launch("k_diag_edge");
}

// Diagnostics of the state in (u, h) into ping-pong slot `slot`
// (compute_solve_diagnostics, plus the divergence the viscosity reads).
int sw_diagnostics(mpas_ctx *ctx, const double *u, const double *h, int slot) {
  SwState &S = *ctx->sw;
  return diagnostics_into(ctx, u, h, S.ke[slot], S.div[slot], S.vort[slot], S.pv_vertex,
                          S.h_edge[slot], S.pv_edge[slot]);
}

// nsteps of apps/sw/run.cc:13-70.
int sw_rk4_steps(mpas_ctx *ctx, double dt, int nsteps, double *const *frozen_rfac) {
  SwState &W = *ctx->sw;
  for (int step = 0; step < nsteps; step++) {
    for (int s = 0; s < 4; s++) {
      const int cur = W.cur, nxt = 1 - W.cur;
      StageIO S = stage_params(ctx, s, dt);
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
      S.rfac = frozen_rfac ? frozen_rfac[4 * step + s] : W.rfac;
      S.keep_rfac = frozen_rfac != nullptr;
      S.tlo = W.tlo;
      if (int rc = launch_stage(ctx, S, W.vort[nxt], W.pv_vertex, W.h_edge[nxt], W.pv_edge[nxt]))
        return rc;
      W.cur = nxt;
    }
    W.elapsed += dt;
  }
  return 0;
}

}  // namespace mpasx
