// ORACLE — test infrastructure only (see orc_internal.h).
// Shallow-water task bodies restated from apps/sw/tasks/solve.cc with the
// same loop orders, scatter orders and expression association.
#include <cmath>

#include "orc_internal.h"

namespace orc {

// apps/sw/tasks/solve.cc:11-107
void sw_compute_solve_diagnostics(const Mesh &m, int nL, const double *h,
                                  const double *u, double *h_edge,
                                  double *h_vertex, double *circulation,
                                  double *vorticity, double *ke,
                                  double *pv_edge, double *pv_vertex) {
  ORC_PAR
  for (int e = 0; e < m.nE; e++) {  // 31-39
    int cell1 = m.e2c.at(e, 0), cell2 = m.e2c.at(e, 1);
    for (int k = 0; k < nL; k++)
      h_edge[e * nL + k] = 0.5 * (h[cell1 * nL + k] + h[cell2 * nL + k]);
  }
  ORC_PAR
  for (int v = 0; v < m.nV; v++)  // 42-46
    for (int k = 0; k < nL; k++) circulation[v * nL + k] = 0.0;
  for (int e = 0; e < m.nE; e++) {  // 47-55
    int v1 = m.e2v.at(e, 0), v2 = m.e2v.at(e, 1);
    for (int k = 0; k < nL; k++) {
      circulation[v1 * nL + k] = circulation[v1 * nL + k] - m.dcEdge[e] * u[e * nL + k];
      circulation[v2 * nL + k] = circulation[v2 * nL + k] + m.dcEdge[e] * u[e * nL + k];
    }
  }
  ORC_PAR
  for (int v = 0; v < m.nV; v++)  // 56-60
    for (int k = 0; k < nL; k++)
      vorticity[v * nL + k] = circulation[v * nL + k] / m.areaTriangle[v];

  ORC_PAR
  for (int c = 0; c < m.nC; c++) {  // 63-77
    for (int k = 0; k < nL; k++) ke[c * nL + k] = 0.0;
    for (int j = 0; j < m.c2e.size(c); j++) {
      int e = m.c2e.at(c, j);
      for (int k = 0; k < nL; k++)
        ke[c * nL + k] += 0.25 * m.dcEdge[e] * m.dvEdge[e] * std::pow(u[e * nL + k], 2);
    }
    for (int k = 0; k < nL; k++) ke[c * nL + k] /= m.areaCell[c];
  }

  ORC_PAR
  for (int v = 0; v < m.nV; v++) {  // 81-92 (ascending cells paired with
    for (int k = 0; k < nL; k++) {  //   file-order kite areas, line 86)
      h_vertex[v * nL + k] = 0.0;
      for (int i = 0; i < 3; i++)
        h_vertex[v * nL + k] += h[m.v2c.at(v, i) * nL + k] * m.kite[v * 3 + i];
      h_vertex[v * nL + k] /= m.areaTriangle[v];
      pv_vertex[v * nL + k] = (m.fVertex[v] + vorticity[v * nL + k]) / h_vertex[v * nL + k];
    }
  }

  ORC_PAR
  for (int e = 0; e < m.nE; e++)  // 95-99
    for (int k = 0; k < nL; k++) pv_edge[e * nL + k] = 0.0;
  for (int v = 0; v < m.nV; v++)  // 100-106
    for (int j = 0; j < m.v2e.size(v); j++) {
      int e = m.v2e.at(v, j);
      for (int k = 0; k < nL; k++) pv_edge[e * nL + k] += 0.5 * pv_vertex[v * nL + k];
    }
}

// apps/sw/tasks/solve.cc:110-141
void sw_init_provis(const Mesh &m, int nL, int nT, const double *u_old,
                    const double *h_old, const double *tr_old,
                    const double *h_edge_old, const double *ke_old,
                    const double *pv_edge_old, double *u_p, double *h_p,
                    double *tr_p, double *h_edge_p, double *ke_p,
                    double *pv_edge_p) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < nL; j++) {
      h_p[c * nL + j] = h_old[c * nL + j];
      ke_p[c * nL + j] = ke_old[c * nL + j];
      for (int k = 0; k < nT; k++)
        tr_p[(c * nL + j) * nT + k] = tr_old[(c * nL + j) * nT + k];
    }
  ORC_PAR
  for (int e = 0; e < m.nE; e++)
    for (int j = 0; j < nL; j++) {
      u_p[e * nL + j] = u_old[e * nL + j];
      h_edge_p[e * nL + j] = h_edge_old[e * nL + j];
      pv_edge_p[e * nL + j] = pv_edge_old[e * nL + j];
    }
}

// apps/sw/tasks/solve.cc:144-166
void sw_setup_timestep(const Mesh &m, int nL, int nT, const double *u_old,
                       const double *h_old, const double *tr_old, double *u_new,
                       double *h_new, double *tr_new) {
  ORC_PAR
  for (int e = 0; e < m.nE; e++)
    for (int j = 0; j < nL; j++) u_new[e * nL + j] = u_old[e * nL + j];
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < nL; j++) {
      h_new[c * nL + j] = h_old[c * nL + j];
      for (int k = 0; k < nT; k++)
        tr_new[(c * nL + j) * nT + k] = tr_old[(c * nL + j) * nT + k] * h_old[c * nL + j];
    }
}

// apps/sw/tasks/solve.cc:169-234
void sw_compute_tend(const Mesh &m, int nL, double gravity,
                     const double *h_edge, const double *h, const double *h_s,
                     const double *u, const double *ke, const double *pv_edge,
                     double *tend_h, double *tend_u) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int k = 0; k < nL; k++) tend_h[c * nL + k] = 0;
  for (int e = 0; e < m.nE; e++) {  // 191-201
    int cell1 = m.e2c.at(e, 0), cell2 = m.e2c.at(e, 1);
    for (int k = 0; k < nL; k++) {
      double flux = u[e * nL + k] * m.dvEdge[e] * h_edge[e * nL + k];
      tend_h[cell1 * nL + k] -= flux;
      tend_h[cell2 * nL + k] += flux;
    }
  }
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int k = 0; k < nL; k++) tend_h[c * nL + k] /= m.areaCell[c];

  ORC_PAR
  for (int e = 0; e < m.nE; e++)
    for (int k = 0; k < nL; k++) tend_u[e * nL + k] = 0.0;
  ORC_PAR
  for (int e = 0; e < m.nE; e++) {  // 215-233
    int cell1 = m.e2c.at(e, 0), cell2 = m.e2c.at(e, 1);
    for (int k = 0; k < nL; k++) {
      double q = 0;
      int j = 0;
      for (int jj = 0; jj < m.e2e.size(e); jj++) {
        int eoe = m.e2e.at(e, jj);
        double workpv = 0.5 * (pv_edge[e * nL + k] + pv_edge[eoe * nL + k]);
        q = q + m.weightsOnEdge[e * 16 + j] * u[eoe * nL + k] * workpv * h_edge[eoe * nL + k];
        ++j;
      }
      tend_u[e * nL + k] =
          q - (ke[cell2 * nL + k] - ke[cell1 * nL + k] +
               gravity * (h[cell2 * nL + k] + h_s[cell2] - h[cell1 * nL + k] - h_s[cell1])) /
                  m.dcEdge[e];
    }
  }
}

// apps/sw/tasks/solve.cc:237-270
void sw_compute_scalar_tend(const Mesh &m, int nL, int nT,
                            const double *h_edge, const double *u,
                            const double *tracers, double *tracer_tend) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int k = 0; k < nL; k++)
      for (int i = 0; i < nT; i++) tracer_tend[(c * nL + k) * nT + i] = 0.0;
  for (int e = 0; e < m.nE; e++) {
    int cell1 = m.e2c.at(e, 0), cell2 = m.e2c.at(e, 1);
    for (int k = 0; k < nL; k++)
      for (int i = 0; i < nT; i++) {
        double tracer_edge =
            0.5 * (tracers[(cell1 * nL + k) * nT + i] + tracers[(cell2 * nL + k) * nT + i]);
        double flux = u[e * nL + k] * m.dvEdge[e] * h_edge[e * nL + k] * tracer_edge;
        tracer_tend[(cell1 * nL + k) * nT + i] -= flux / m.areaCell[cell1];
        tracer_tend[(cell2 * nL + k) * nT + i] += flux / m.areaCell[cell2];
      }
  }
}

// apps/sw/tasks/solve.cc:273-314
void sw_compute_substep(const Mesh &m, int nL, int nT, int freeze_u,
                        double *u_p, const double *u_old, const double *u_tend,
                        double *h_p, const double *h_old, const double *h_tend,
                        double *tr_p, const double *tr_old,
                        const double *tr_tend, double rkweight) {
  ORC_PAR
  for (int e = 0; e < m.nE; e++)
    for (int j = 0; j < nL; j++)
      u_p[e * nL + j] = u_old[e * nL + j] + rkweight * u_tend[e * nL + j];
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < nL; j++)
      h_p[c * nL + j] = h_old[c * nL + j] + rkweight * h_tend[c * nL + j];
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < nL; j++)
      for (int k = 0; k < nT; k++)
        tr_p[(c * nL + j) * nT + k] =
            (h_old[c * nL + j] * tr_old[(c * nL + j) * nT + k] +
             rkweight * tr_tend[(c * nL + j) * nT + k]) /
            h_p[c * nL + j];
  if (freeze_u)
    for (int e = 0; e < m.nE; e++)
      for (int j = 0; j < nL; j++) u_p[e * nL + j] = u_old[e * nL + j];
}

// apps/sw/tasks/solve.cc:317-341
void sw_accumulate_update(const Mesh &m, int nL, int nT, double *u_new,
                          const double *u_tend, double *h_new,
                          const double *h_tend, double *tr_new,
                          const double *tr_tend, double rkweight) {
  ORC_PAR
  for (int e = 0; e < m.nE; e++)
    for (int j = 0; j < nL; j++) u_new[e * nL + j] += rkweight * u_tend[e * nL + j];
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < nL; j++) {
      h_new[c * nL + j] += rkweight * h_tend[c * nL + j];
      for (int k = 0; k < nT; k++)
        tr_new[(c * nL + j) * nT + k] += rkweight * tr_tend[(c * nL + j) * nT + k];
    }
}

// apps/sw/tasks/solve.cc:344-367
void sw_finalize_timestep(const Mesh &m, int nL, int nT, int freeze_u,
                          const double *u_old, double *u_new,
                          const double *h_new, double *tr_new) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < nL; j++)
      for (int k = 0; k < nT; k++) tr_new[(c * nL + j) * nT + k] /= h_new[c * nL + j];
  if (freeze_u)
    for (int e = 0; e < m.nE; e++)
      for (int j = 0; j < nL; j++) u_new[e * nL + j] = u_old[e * nL + j];
}

// apps/sw/tasks/solve.cc:370-401
void sw_timeshift(const Mesh &m, int nL, int nT, double *u_old, double *h_old,
                  double *tr_old, double *h_edge_old, double *ke_old,
                  double *pv_edge_old, const double *u_new, const double *h_new,
                  const double *tr_new, const double *h_edge_new,
                  const double *ke_new, const double *pv_edge_new) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < nL; j++) {
      h_old[c * nL + j] = h_new[c * nL + j];
      ke_old[c * nL + j] = ke_new[c * nL + j];
      for (int k = 0; k < nT; k++)
        tr_old[(c * nL + j) * nT + k] = tr_new[(c * nL + j) * nT + k];
    }
  ORC_PAR
  for (int e = 0; e < m.nE; e++)
    for (int j = 0; j < nL; j++) {
      u_old[e * nL + j] = u_new[e * nL + j];
      h_edge_old[e * nL + j] = h_edge_new[e * nL + j];
      pv_edge_old[e * nL + j] = pv_edge_new[e * nL + j];
    }
}

// mpasoflecsi/eqns/metrics.hh:14-21
static double sphere_distance(double lat1, double lon1, double lat2, double lon2,
                              double radius) {
  auto arg1 = std::sqrt(std::pow(std::sin(0.5 * (lat2 - lat1)), 2) +
                        std::cos(lat1) * std::cos(lat2) *
                            std::pow(std::sin(0.5 * (lon2 - lon1)), 2));
  return 2 * radius * std::asin(arg1);
}

// apps/sw/tasks/solve.cc:404-436 (case-1 L2 error of h, level 0)
double sw_compute_error(const Mesh &m, int nL, double a, const double *h,
                        double elapsed) {
  const double pi = 3.141592653589793;
  constexpr double h0 = 1000.0;
  auto day_frac = elapsed / 86400 / 12;
  auto rot_ang = (3 * pi / 2.) + 2 * pi * day_frac;
  double num = 0, den = 0;
  for (int c = 0; c < m.nC; c++) {
    double h_true;
    auto r = sphere_distance(0.0, rot_ang, m.latCell[c], m.lonCell[c], a);
    if (r < a / 3.0)
      h_true = (h0 / 2.0) * (1.0 + std::cos(pi * r * 3.0 / a));
    else
      h_true = h0 / 2.0;
    num += std::pow(h[c * nL] - h_true, 2) * std::cos(m.latCell[c]) / m.areaCell[c];
    den += std::pow(h_true, 2) * std::cos(m.latCell[c]) / m.areaCell[c];
  }
  num /= (4 * pi);
  den /= (4 * pi);
  return std::sqrt(num) / std::sqrt(den);
}

// Del2 viscosity on the normal velocity.  Not in the reference (its
// compute_tend has no dissipation, apps/sw/tasks/solve.cc:228-231); defined
// here after MPAS-Ocean's del2 momentum mixing:
//   div(c)  = (sum_{edgesOnCell, file order} (-edgeSign) * dvEdge * u) / areaCell
//   tend_u += (meshScalingDel2 * visc2) *
//             ((div(c2) - div(c1)) / dcEdge - (vort(v2) - vort(v1)) / dvEdge)
// with the vorticity of compute_solve_diagnostics.
void sw_del2_viscosity(const Mesh &m, int nL, double visc2, const double *u,
                       const double *vorticity, double *divergence,
                       double *tend_u) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++) {
    for (int k = 0; k < nL; k++) divergence[c * nL + k] = 0.0;
    for (int j = 0; j < m.c2e.size(c); j++) {
      int e = m.c2e.at(c, j);
      double s = m.edgeSignOnCell[c * 8 + j];
      for (int k = 0; k < nL; k++)
        divergence[c * nL + k] += (-s) * m.dvEdge[e] * u[e * nL + k];
    }
    for (int k = 0; k < nL; k++) divergence[c * nL + k] /= m.areaCell[c];
  }
  ORC_PAR
  for (int e = 0; e < m.nE; e++) {
    int c1 = m.e2c.at(e, 0), c2 = m.e2c.at(e, 1);
    int v1 = m.e2v.at(e, 0), v2 = m.e2v.at(e, 1);
    double coef = m.meshScalingDel2[e] * visc2;
    for (int k = 0; k < nL; k++) {
      double ud = (divergence[c2 * nL + k] - divergence[c1 * nL + k]) / m.dcEdge[e] -
                  (vorticity[v2 * nL + k] - vorticity[v1 * nL + k]) / m.dvEdge[e];
      tend_u[e * nL + k] += coef * ud;
    }
  }
}

}  // namespace orc
