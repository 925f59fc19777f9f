// Per-entity operators shared by the per-task kernels and the stage-fused
// RK4 kernels, so both paths form every value with the same operations.
#pragma once
#include "dmesh.h"

namespace mpasx {

// compute_solve_diagnostics, vertex part (apps/sw/tasks/solve.cc:42-60,
// 81-92): circulation gathers v1 -= dc*u / v2 += dc*u over the vertex's
// edges in ascending id; h_vertex pairs the ascending m.cells(v) with the
// file-order kiteAreasOnVertex (line 86).  circulation / h_vertex may be
// null (the fused step does not keep them).  Levels [lv0, lv1) (all by
// default), LP at a time (every load of a group before its stores).
template <int LP = 1>
__device__ __forceinline__ void diag_vertex(const DMesh &M, int v, const double *__restrict__ h,
                                            const double *__restrict__ u, double *h_vertex,
                                            double *circulation, double *vorticity,
                                            double *pv_vertex, uint32_t lv0 = 0,
                                            uint32_t lv1 = ~0u) {
  const uint32_t L = M.nL, V = M.nV;
  const int x0 = ldc(M.v2e, v), x1 = ldc(M.v2e, V + v), x2 = ldc(M.v2e, 2 * V + v);
  const uint32_t e0 = v2e_id(x0), e1 = v2e_id(x1), e2 = v2e_id(x2);
  const uint32_t c0 = ldc(M.v2c, v), c1 = ldc(M.v2c, V + v), c2 = ldc(M.v2c, 2 * V + v);
  const double k0 = ldc(M.kite, v), k1 = ldc(M.kite, V + v), k2 = ldc(M.kite, 2 * V + v);
  const double d0 = ldc(M.dc, e0), d1 = ldc(M.dc, e1), d2 = ldc(M.dc, e2);
  const double at = ldc(M.areaTri, v), f = ldc(M.fV, v);
  for (uint32_t k = lv0; k < min(lv1, L); k += LP) {
    double u0[LP], u1[LP], u2[LP], h0[LP], h1[LP], h2[LP];
    ld_lv<LP>(u, e0 * L + k, u0);
    ld_lv<LP>(u, e1 * L + k, u1);
    ld_lv<LP>(u, e2 * L + k, u2);
    ld_lv<LP>(h, c0 * L + k, h0);
    ld_lv<LP>(h, c1 * L + k, h1);
    ld_lv<LP>(h, c2 * L + k, h2);
    double circ[LP], vort[LP], hv[LP], pvv[LP];
#pragma unroll
    for (int p = 0; p < LP; p++) {
      double c = 0.0;
      c = c + v2e_sign(x0) * (d0 * u0[p]);
      c = c + v2e_sign(x1) * (d1 * u1[p]);
      c = c + v2e_sign(x2) * (d2 * u2[p]);
      circ[p] = c;
      vort[p] = c / at;
      double s = 0.0;
      s += h0[p] * k0;
      s += h1[p] * k1;
      s += h2[p] * k2;
      hv[p] = s / at;
      pvv[p] = (f + vort[p]) / hv[p];
    }
    const uint32_t o = v * L + k;
    if (circulation) st_lv<LP>(circulation, o, circ);
    st_lv<LP>(vorticity, o, vort);
    if (h_vertex) st_lv<LP>(h_vertex, o, hv);
    st_lv<LP>(pv_vertex, o, pvv);
  }
}

// compute_solve_diagnostics, edge part (solve.cc:31-39, 95-106): pv_edge
// starts at zero and takes half of each vertex value, lower vertex id first.
template <int LP = 1>
__device__ __forceinline__ void diag_edge(const DMesh &M, int e, const double *__restrict__ h,
                                          const double *__restrict__ pv_vertex,
                                          double *h_edge, double *pv_edge, uint32_t lv0 = 0,
                                          uint32_t lv1 = ~0u) {
  const uint32_t L = M.nL;
  const int2 cc = ldc(M.e2c, e);
  const int2 vv = ldc(M.e2v, e);
  const uint32_t va = min(vv.x, vv.y), vb = max(vv.x, vv.y);
  for (uint32_t k = lv0; k < min(lv1, L); k += LP) {
    const uint32_t o = e * L + k;
    double h1[LP], h2[LP], pa[LP], pb[LP];
    ld_lv<LP>(h, uint32_t(cc.x) * L + k, h1);
    ld_lv<LP>(h, uint32_t(cc.y) * L + k, h2);
    ld_lv<LP>(pv_vertex, va * L + k, pa);
    ld_lv<LP>(pv_vertex, vb * L + k, pb);
    double he[LP], pe[LP];
#pragma unroll
    for (int p = 0; p < LP; p++) {
      he[p] = 0.5 * (h1[p] + h2[p]);
      double q = 0.0;
      q += 0.5 * pa[p];
      q += 0.5 * pb[p];
      pe[p] = q;
    }
    st_lv<LP>(h_edge, o, he);
    st_lv<LP>(pv_edge, o, pe);
  }
}

// compute_tend, normal-velocity part (solve.cc:215-233).
__device__ __forceinline__ double tend_u_at(const DMesh &M, int e, int k, int2 cc,
                                            const double *__restrict__ u,
                                            const double *__restrict__ h_edge,
                                            const double *__restrict__ pv_edge,
                                            const double *__restrict__ ke,
                                            const double *__restrict__ h,
                                            const double *__restrict__ h_s) {
  const uint32_t L = M.nL, E = M.nE;
  const double pve = ld(pv_edge, e * L + k);
  const int n = ldc(M.nEoE, e);
  double q = 0;
  for (int j = 0; j < n; j++) {
    const uint32_t eoe = ldc(M.eoe, j * E + e);
    const double workpv = 0.5 * (pve + ld(pv_edge, eoe * L + k));
    q = q + ldc(M.woe, j * E + e) * ld(u, eoe * L + k) * workpv * ld(h_edge, eoe * L + k);
  }
  const uint32_t o1 = uint32_t(cc.x) * L + k, o2 = uint32_t(cc.y) * L + k;
  return q - (ld(ke, o2) - ld(ke, o1) +
              M.gravity * (ld(h, o2) + ld(h_s, cc.y) - ld(h, o1) - ld(h_s, cc.x))) /
                 ldc(M.dc, e);
}

// del2 viscosity on u (no reference counterpart):
// coef * ((div(c2) - div(c1)) / dcEdge - (vort(v2) - vort(v1)) / dvEdge).
__device__ __forceinline__ double visc_at(const DMesh &M, int e, int k, int2 cc, int2 vv,
                                          const double *__restrict__ div,
                                          const double *__restrict__ vort, double coef) {
  const uint32_t L = M.nL;
  const double ud = (ld(div, uint32_t(cc.y) * L + k) - ld(div, uint32_t(cc.x) * L + k)) /
                        ldc(M.dc, e) -
                    (ld(vort, uint32_t(vv.y) * L + k) - ld(vort, uint32_t(vv.x) * L + k)) /
                        ldc(M.dv, e);
  return coef * ud;
}

}  // namespace mpasx
