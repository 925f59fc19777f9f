// Per-entity operators shared by the per-task kernels and the stage-fused
// RK4 kernels, so both paths form every value with the same operations.
#pragma once
#include "dmesh.h"

namespace mpasx {

// compute_solve_diagnostics, vertex part (apps/sw/tasks/solve.cc:42-60,
// 81-92): circulation gathers v1 -= dc*u / v2 += dc*u over the vertex's
// edges in ascending id; h_vertex pairs the ascending m.cells(v) with the
// file-order kiteAreasOnVertex (line 86).  circulation / h_vertex may be
// null (the fused step does not keep them).  Levels [lv0, lv1) (all by default).
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
  for (uint32_t k = lv0; k < min(lv1, L); k++) {
    double circ = 0.0;
    circ = circ + v2e_sign(x0) * (d0 * ld(u, e0 * L + k));
    circ = circ + v2e_sign(x1) * (d1 * ld(u, e1 * L + k));
    circ = circ + v2e_sign(x2) * (d2 * ld(u, e2 * L + k));
    const double vort = circ / at;
    double hv = 0.0;
    hv += ld(h, c0 * L + k) * k0;
    hv += ld(h, c1 * L + k) * k1;
    hv += ld(h, c2 * L + k) * k2;
    hv /= at;
    const uint32_t o = v * L + k;
    if (circulation) st(circulation, o, circ);
    st(vorticity, o, vort);
    if (h_vertex) st(h_vertex, o, hv);
    st(pv_vertex, o, (f + vort) / hv);
  }
}

// compute_solve_diagnostics, edge part (solve.cc:31-39, 95-106): pv_edge
// starts at zero and takes half of each vertex value, lower vertex id first.
__device__ __forceinline__ void diag_edge(const DMesh &M, int e, const double *__restrict__ h,
                                          const double *__restrict__ pv_vertex,
                                          double *h_edge, double *pv_edge, uint32_t lv0 = 0,
                                          uint32_t lv1 = ~0u) {
  const uint32_t L = M.nL;
  const int2 cc = ldc(M.e2c, e);
  const int2 vv = ldc(M.e2v, e);
  const uint32_t va = min(vv.x, vv.y), vb = max(vv.x, vv.y);
  for (uint32_t k = lv0; k < min(lv1, L); k++) {
    const uint32_t o = e * L + k;
    st(h_edge, o, 0.5 * (ld(h, uint32_t(cc.x) * L + k) + ld(h, uint32_t(cc.y) * L + k)));
    double p = 0.0;
    p += 0.5 * ld(pv_vertex, va * L + k);
    p += 0.5 * ld(pv_vertex, vb * L + k);
    st(pv_edge, o, p);
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
