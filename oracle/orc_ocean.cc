// ORACLE — test infrastructure only (see orc_internal.h).
// FCT tracer advection and del2 tracer mixing restated from
// apps/ocean/tasks/solve.cc with the same loop orders and association.
#include <algorithm>
#include <vector>

#include "orc_internal.h"

namespace orc {

#define TR(c, k, t) (((c) * nL + (k)) * nT + (t))

// apps/ocean/tasks/solve.cc:13-26
void ocn_copy_tracers(const Mesh &m, int nL, int nT, const double *tr,
                      double *trNew, double *trPrev) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int k = 0; k < nL; k++)
      for (int t = 0; t < nT; t++) {
        trNew[TR(c, k, t)] = tr[TR(c, k, t)];
        trPrev[TR(c, k, t)] = tr[TR(c, k, t)];
      }
}

// apps/ocean/tasks/solve.cc:28-38
void ocn_reinitialize_tracer_tendencies(const Mesh &m, int nL, int nT, double *tend) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int k = 0; k < nL; k++)
      for (int t = 0; t < nT; t++) tend[TR(c, k, t)] = 0.0;
}

// apps/ocean/tasks/solve.cc:40-57
void ocn_calc_normal_thickness_flux(const Mesh &m, int nL,
                                    const double *layerThicknessEdge,
                                    const double *normalVelocity,
                                    double *normalThicknessFlux) {
  ORC_PAR
  for (int e = 0; e < m.nE; e++)
    for (int k = 0; k < nL; k++)
      normalThicknessFlux[e * nL + k] = normalVelocity[e * nL + k] * layerThicknessEdge[e * nL + k];
}

// apps/ocean/tasks/solve.cc:60-97
void ocn_calc_provisional_layer_thicknesses(const Mesh &m, int nL,
                                            const double *layerThickness,
                                            const double *normalThicknessFlux,
                                            double *hProv, double *hProvInv,
                                            double *hNewInv, double dt) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++) {
    double dtInvAreaCell = dt / m.areaCell[c];
    int kmax = m.maxLevelCell[c];
    for (int k = 0; k < kmax; k++) hProv[c * nL + k] = layerThickness[c * nL + k];
    int je = 0;
    for (int j = 0; j < m.c2e.size(c); j++) {
      int e = m.c2e.at(c, j);
      double signedFactor = dtInvAreaCell * m.dvEdge[e] * m.edgeSignOnCell[c * 8 + je];
      for (int k = 0; k < kmax; k++)
        hProv[c * nL + k] += signedFactor * normalThicknessFlux[e * nL + k];
      je++;
    }
    for (int k = 0; k < kmax; k++) {
      hProvInv[c * nL + k] = 1.0 / hProv[c * nL + k];
      hNewInv[c * nL + k] = 1.0 / (hProv[c * nL + k]);
    }
  }
}

// apps/ocean/tasks/solve.cc:99-118
void ocn_low_order(const Mesh &m, int nL, int nT,
                   const double *normalThicknessFlux, const double *tracers,
                   double *lowOrderFlux) {
  ORC_PAR
  for (int e = 0; e < m.nE; e++) {
    int c1 = m.e2c.at(e, 0), c2 = m.e2c.at(e, 1);
    for (int k = 0; k < nL; k++) {
      double ntf1 = std::max(0.0, normalThicknessFlux[e * nL + k]);
      double ntf2 = std::min(0.0, normalThicknessFlux[e * nL + k]);
      for (int t = 0; t < nT; t++)
        lowOrderFlux[(e * nL + k) * nT + t] =
            m.dvEdge[e] * (ntf1 * tracers[TR(c1, k, t)] + ntf2 * tracers[TR(c2, k, t)]);
    }
  }
}

// apps/ocean/tasks/solve.cc:120-167
void ocn_high_order(const Mesh &m, int nL, int nT, double coef3rdOrder,
                    const double *normalThicknessFlux, const double *tracers,
                    double *highOrderFlux) {
#pragma omp parallel
  {
  std::vector<double> wgtTmp(nL), sgnTmp(nL);
  std::vector<std::vector<double>> flxTmp(nL, std::vector<double>(nT));
#pragma omp for schedule(static)
  for (int e = 0; e < m.nE; e++) {
    for (int k = 0; k < nL; k++) {
      wgtTmp[k] = normalThicknessFlux[e * nL + k] * m.highOrderAdvMask[e * nL + k];
      sgnTmp[k] = (normalThicknessFlux[e * nL + k] < 0.0) ? -1.0
                                                          : (normalThicknessFlux[e * nL + k] > 0.);
      for (int t = 0; t < nT; t++) flxTmp[k][t] = 0.0;
    }
    for (int i = 0; i < m.nAdvCellsForEdge[e]; i++) {
      int iCell = m.advCellsForEdge[e * 16 + i];
      double coef1 = m.advCoef[e * 16 + i];
      double coef3 = m.advCoef_3rd[e * 16 + i] * coef3rdOrder;
      for (int k = 0; k < m.maxLevelCell[iCell]; k++) {
        double weightFactor = wgtTmp[k] * (coef1 + coef3 * sgnTmp[k]);
        for (int t = 0; t < nT; t++) flxTmp[k][t] += tracers[TR(iCell, k, t)] * weightFactor;
      }
    }
    for (int k = 0; k < nL; k++)
      for (int t = 0; t < nT; t++) highOrderFlux[(e * nL + k) * nT + t] = flxTmp[k][t];
  }
  }
}

// apps/ocean/tasks/solve.cc:169-285
void ocn_correct_and_rescale(const Mesh &m, int nL, int nT,
                             const double *hProvInv,
                             const double *layerThickness,
                             const double *normalThicknessFlux,
                             const double *lowOrderFlux, const double *tracers,
                             double *highOrderFlux, double *flxIn,
                             double *flxOut, double *tend, double *tracerMax,
                             double *tracerMin, double dt) {
  const double eps = 1.e-10;
#define EF(e, k, t) (((e) * nL + (k)) * nT + (t))
  ORC_PAR
  for (int c = 0; c < m.nC; c++) {  // 195-214
    for (int k = 0; k < m.maxLevelCell[c]; k++)
      for (int t = 0; t < nT; t++) {
        tracerMax[TR(c, k, t)] = tracers[TR(c, k, t)];
        tracerMin[TR(c, k, t)] = tracers[TR(c, k, t)];
        flxIn[TR(c, k, t)] = 0.0;
        flxOut[TR(c, k, t)] = 0.0;
      }
    for (int j = 0; j < m.c2c.size(c); j++) {
      int c2 = m.c2c.at(c, j);
      int kmax = std::min(m.maxLevelCell[c], m.maxLevelCell[c2]);
      for (int k = 0; k < kmax; k++)
        for (int t = 0; t < nT; t++) {
          tracerMax[TR(c, k, t)] = std::max(tracerMax[TR(c, k, t)], tracers[TR(c2, k, t)]);
          tracerMin[TR(c, k, t)] = std::min(tracerMin[TR(c, k, t)], tracers[TR(c2, k, t)]);
        }
    }
  }
  ORC_PAR
  for (int e = 0; e < m.nE; e++) {  // 216-229
    int c1 = m.e2c.at(e, 0), c2 = m.e2c.at(e, 1);
    for (int k = 0; k < m.maxLevelEdgeTop[e]; k++) {
      double tracerWeight = (m.highOrderAdvMask[e * nL + k] + 1 & 1) * (m.dvEdge[e] * 0.5) *
                            normalThicknessFlux[e * nL + k];
      for (int t = 0; t < nT; t++) {
        highOrderFlux[EF(e, k, t)] += tracerWeight * (tracers[TR(c1, k, t)] + tracers[TR(c2, k, t)]);
        highOrderFlux[EF(e, k, t)] -= lowOrderFlux[EF(e, k, t)];
      }
    }
  }
  ORC_PAR
  for (int c = 0; c < m.nC; c++) {  // 232-267
    double invAreaCell1 = 1.0 / m.areaCell[c];
    int je = 0;
    for (int j = 0; j < m.c2e.size(c); j++) {
      int e = m.c2e.at(c, j);
      double signedFactor = m.edgeSignOnCell[c * 8 + je] * invAreaCell1;
      for (int k = 0; k < m.maxLevelEdgeTop[e]; k++)
        for (int t = 0; t < nT; t++) {
          tend[TR(c, k, t)] += signedFactor * lowOrderFlux[EF(e, k, t)];
          flxIn[TR(c, k, t)] += std::max(0.0, signedFactor * highOrderFlux[EF(e, k, t)]);
          flxOut[TR(c, k, t)] += std::min(0.0, signedFactor * highOrderFlux[EF(e, k, t)]);
        }
      je++;
    }
    for (int k = 0; k < m.maxLevelCell[c]; k++)
      for (int t = 0; t < nT; t++) {
        double tracerUpwindNew =
            (tracers[TR(c, k, t)] * layerThickness[c * nL + k] + dt * tend[TR(c, k, t)]) *
            hProvInv[c * nL + k];
        double tracerMaxNew = tracerUpwindNew + dt * flxIn[TR(c, k, t)] * hProvInv[c * nL + k];
        double tracerMinNew = tracerUpwindNew + dt * flxOut[TR(c, k, t)] * hProvInv[c * nL + k];
        double scaleFactor = (tracerMax[TR(c, k, t)] - tracerUpwindNew) /
                             (tracerMaxNew - tracerUpwindNew + eps);
        flxIn[TR(c, k, t)] = std::min(1.0, std::max(0.0, scaleFactor));
        scaleFactor = (tracerUpwindNew - tracerMin[TR(c, k, t)]) /
                      (tracerUpwindNew - tracerMinNew + eps);
        flxOut[TR(c, k, t)] = std::min(1.0, std::max(0.0, scaleFactor));
      }
  }
  ORC_PAR
  for (int e = 0; e < m.nE; e++) {  // 270-282
    int c1 = m.e2c.at(e, 0), c2 = m.e2c.at(e, 1);
    for (int k = 0; k < m.maxLevelEdgeTop[e]; k++)
      for (int t = 0; t < nT; t++)
        highOrderFlux[EF(e, k, t)] =
            std::max(0.0, highOrderFlux[EF(e, k, t)]) *
                std::min(flxOut[TR(c1, k, t)], flxIn[TR(c2, k, t)]) +
            std::min(0.0, highOrderFlux[EF(e, k, t)]) *
                std::min(flxIn[TR(c1, k, t)], flxOut[TR(c2, k, t)]);
  }
#undef EF
}

// apps/ocean/tasks/solve.cc:287-326
void ocn_accumulate_fct(const Mesh &m, int nL, int nT, const double *hProvInv,
                        const double *layerThickness,
                        const double *highOrderFlux, const double *tracers,
                        double *tracersProv, double *tend, double dt) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++) {
    double invAreaCell1 = 1.0 / m.areaCell[c];
    int je = 0;
    for (int j = 0; j < m.c2e.size(c); j++) {
      int e = m.c2e.at(c, j);
      double signedFactor = m.edgeSignOnCell[c * 8 + je] * invAreaCell1;
      for (int k = 0; k < m.maxLevelEdgeTop[e]; k++)
        for (int t = 0; t < nT; t++)
          tend[TR(c, k, t)] += signedFactor * highOrderFlux[((e * nL) + k) * nT + t];
      je++;
    }
    for (int k = 0; k < m.maxLevelCell[c]; k++)
      for (int t = 0; t < nT; t++)
        tracersProv[TR(c, k, t)] =
            (tracers[TR(c, k, t)] * layerThickness[c * nL + k] + dt * tend[TR(c, k, t)]) *
            hProvInv[c * nL + k];
  }
}

// apps/ocean/tasks/solve.cc:328-366
void ocn_hmix_del2(const Mesh &m, int nL, int nT,
                   const double *layerThicknessEdge, const double *tracers,
                   double *tend, double coef) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++) {
    double invArea = 1.0 / m.areaCell[c];
    int je = 0;
    for (int j = 0; j < m.c2e.size(c); j++) {
      int e = m.c2e.at(c, j);
      int c1 = m.e2c.at(e, 0), c2 = m.e2c.at(e, 1);
      double rho_scaling = m.meshScalingDel2[e] * coef * m.dvEdge[e] / m.dcEdge[e];
      for (int k = 0; k < m.maxLevelEdgeTop[e]; k++)
        for (int t = 0; t < nT; t++) {
          double tracer_turb_flux = tracers[TR(c2, k, t)] - tracers[TR(c1, k, t)];
          double flux = layerThicknessEdge[e * nL + k] * tracer_turb_flux * rho_scaling;
          tend[TR(c, k, t)] -= m.edgeSignOnCell[c * 8 + je] * flux * invArea;
        }
      je++;
    }
  }
}

// apps/ocean/tasks/solve.cc:368-399
void ocn_step_tracers_forward(const Mesh &m, int nL, int nT, const double *tend,
                              const double *tracers, double *tracersNew,
                              double *tracersPrev, double rk_weight,
                              double rk_subweight) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < nL; j++)
      for (int k = 0; k < nT; k++) {
        tracersNew[TR(c, j, k)] += rk_weight * tend[TR(c, j, k)];
        tracersPrev[TR(c, j, k)] = tracers[TR(c, j, k)] + rk_subweight * tend[TR(c, j, k)];
      }
}

// apps/ocean/tasks/solve.cc:401-414
void ocn_update_tracers(const Mesh &m, int nL, int nT, const double *trNew, double *tr) {
  ORC_PAR
  for (int c = 0; c < m.nC; c++)
    for (int k = 0; k < nL; k++)
      for (int t = 0; t < nT; t++) tr[TR(c, k, t)] = trNew[TR(c, k, t)];
}

#undef TR
}  // namespace orc
