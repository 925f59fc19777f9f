// Per-entity operators shared by the per-task kernels and the stage-fused
// RK4 kernels, so both paths form every value with the same operations.
#pragma once
#include "dmesh.h"

namespace mpasx {

// compute_solve_diagnostics, vertex part (apps/sw/tasks/solve.cc:42-60,
// 81-92): circulation gathers v1 -= dc*u / v2 += dc*u over the vertex's
// edges in ascending id; h_vertex pairs the ascending m.cells(v) with the
// file-order kiteAreasOnVertex (line 86).  circulation / h_vertex may be
// null (the fused step does not keep them).
__device__ __forceinline__ void diag_vertex(const DMesh &M, int v, const double *__restrict__ h,
                                            const double *__restrict__ u, double *h_vertex,
                                            double *circulation, double *vorticity,
                                            double *pv_vertex) {
  const int L = M.nL, V = M.nV;
  const int x0 = M.v2e[v], x1 = M.v2e[V + v], x2 = M.v2e[2 * V + v];
  const int e0 = v2e_id(x0), e1 = v2e_id(x1), e2 = v2e_id(x2);
  const int c0 = M.v2c[v], c1 = M.v2c[V + v], c2 = M.v2c[2 * V + v];
  const double k0 = M.kite[v], k1 = M.kite[V + v], k2 = M.kite[2 * V + v];
  const double d0 = M.dc[e0], d1 = M.dc[e1], d2 = M.dc[e2];
  const double at = M.areaTri[v], f = M.fV[v];
  for (int k = 0; k < L; k++) {
    double circ = 0.0;
    circ = circ + v2e_sign(x0) * (d0 * u[size_t(e0) * L + k]);
    circ = circ + v2e_sign(x1) * (d1 * u[size_t(e1) * L + k]);
    circ = circ + v2e_sign(x2) * (d2 * u[size_t(e2) * L + k]);
    const double vort = circ / at;
    double hv = 0.0;
    hv += h[size_t(c0) * L + k] * k0;
    hv += h[size_t(c1) * L + k] * k1;
    hv += h[size_t(c2) * L + k] * k2;
    hv /= at;
    const size_t o = size_t(v) * L + k;
    if (circulation) circulation[o] = circ;
    vorticity[o] = vort;
    if (h_vertex) h_vertex[o] = hv;
    pv_vertex[o] = (f + vort) / hv;
  }
}

// compute_solve_diagnostics, edge part (solve.cc:31-39, 95-106): pv_edge
// starts at zero and takes half of each vertex value, lower vertex id first.
__device__ __forceinline__ void diag_edge(const DMesh &M, int e, const double *__restrict__ h,
                                          const double *__restrict__ pv_vertex,
                                          double *h_edge, double *pv_edge) {
  const int L = M.nL;
  const int2 cc = M.e2c[e];
  const int2 vv = M.e2v[e];
  const int va = min(vv.x, vv.y), vb = max(vv.x, vv.y);
  for (int k = 0; k < L; k++) {
    const size_t o = size_t(e) * L + k;
    h_edge[o] = 0.5 * (h[size_t(cc.x) * L + k] + h[size_t(cc.y) * L + k]);
    double p = 0.0;
    p += 0.5 * pv_vertex[size_t(va) * L + k];
    p += 0.5 * pv_vertex[size_t(vb) * L + k];
    pv_edge[o] = p;
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
  const int L = M.nL;
  const double pve = pv_edge[size_t(e) * L + k];
  const int n = M.nEoE[e];
  double q = 0;
  for (int j = 0; j < n; j++) {
    const int eoe = M.eoe[size_t(j) * M.nE + e];
    const double workpv = 0.5 * (pve + pv_edge[size_t(eoe) * L + k]);
    q = q + M.woe[size_t(j) * M.nE + e] * u[size_t(eoe) * L + k] * workpv *
                h_edge[size_t(eoe) * L + k];
  }
  return q - (ke[size_t(cc.y) * L + k] - ke[size_t(cc.x) * L + k] +
              M.gravity * (h[size_t(cc.y) * L + k] + h_s[cc.y] - h[size_t(cc.x) * L + k] -
                           h_s[cc.x])) /
                 M.dc[e];
}

// del2 viscosity on u (no reference counterpart):
// coef * ((div(c2) - div(c1)) / dcEdge - (vort(v2) - vort(v1)) / dvEdge).
__device__ __forceinline__ double visc_at(const DMesh &M, int e, int k, int2 cc, int2 vv,
                                          const double *__restrict__ div,
                                          const double *__restrict__ vort, double coef) {
  const int L = M.nL;
  const double ud = (div[size_t(cc.y) * L + k] - div[size_t(cc.x) * L + k]) / M.dc[e] -
                    (vort[size_t(vv.y) * L + k] - vort[size_t(vv.x) * L + k]) / M.dv[e];
  return coef * ud;
}

}  // namespace mpasx
