// ORACLE — test infrastructure only.  A CPU restatement of the reference's
// shallow-water and FCT task bodies (apps/sw/tasks/solve.cc,
// apps/ocean/tasks/solve.cc, apps/ocean/tasks/init_functions.cc), its
// drivers (apps/sw/run.cc, apps/ocean/run.cc) and its topology flattening
// (mpasoflecsi/specialization/mesh.cc:67-116, definition.hh:670-718).
// Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use
// it, and only as the checker; the product never links it.
//
// Layouts follow the reference: per-entity arrays of levels (vltensor), and
// [level][tracer] per cell (vltracer), with runtime nL / nT instead of the
// compile-time nVertLevels / maxTracers (mpasoflecsi/common/constants.hh:16-18).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

// Loops whose every iteration writes only its own entity's values (gathers)
// may run on several OpenMP threads: each value is then formed by the same
// operations in the same order as on one thread, so the bits do not change.
// The scatter loops of the reference (several entities adding into one, in
// ascending source order) stay serial.  orc_set_threads picks the count.
#define ORC_PAR _Pragma("omp parallel for schedule(static)")

namespace orc {

// FleCSI ragged connectivity: one list per entity.
struct Crs {
  std::vector<int> off{0};
  std::vector<int> idx;
  int size(int i) const { return off[i + 1] - off[i]; }
  const int *row(int i) const { return idx.data() + off[i]; }
  int at(int i, int j) const { return idx[off[i] + j]; }
  int rows() const { return int(off.size()) - 1; }
};

struct Params {
  int nL = 1;             // vertical levels (reference nVertLevels)
  int nT = 0;             // tracers (reference maxTracers)
  int tracer_adv = 0;     // 0 none, 1 centred (compute_scalar_tend), 2 FCT
  int freeze_u = 0;       // SW test case 1 (apps/sw/tasks/solve.cc:307-313)
  double gravity = 9.80616;
  double a = 6371000.0;   // sphere radius used by setup code
  double coef3rd = 0.25;  // apps/ocean/tasks/solve.cc:11
  double visc2 = 0.0;     // del2 viscosity on u (not in the reference)
  double kappa2 = 0.0;    // del2 tracer diffusion (tracer_hmix_del2_tend coef)
};

struct Mesh {
  int nC = 0, nE = 0, nV = 0;
  int maxEdges = 8, maxEdges2 = 16;
  // connectivity as mesh::initialize builds it
  Crs c2v, c2e, c2c, e2v, e2e, v2c, e2c, v2e;
  // mesh fields (reference layouts: me2tensor = 16, vdtensor = 3)
  std::vector<double> areaCell, dvEdge, dcEdge, areaTriangle, kite, fVertex,
      weightsOnEdge, xCell, yCell, zCell, xVertex, yVertex, zVertex, latCell,
      lonCell, latEdge, lonEdge, latVertex, lonVertex, angleEdge, meshDensity;
  // raw file connectivity kept until finalize
  std::vector<int32_t> f_voc, f_eoc, f_coc, f_voe, f_eoe;

  // FCT / del2 setup products (apps/ocean/tasks/init_functions.cc)
  std::vector<int> maxLevelCell, maxLevelEdgeTop, nAdvCellsForEdge,
      advCellsForEdge /*E*16*/, highOrderAdvMask /*E*nL*/,
      boundaryCell /*C*nL*/;
  std::vector<double> edgeSignOnCell /*C*8*/, meshScalingDel2, derivTwo /*E*16*2*/,
      advCoef /*E*16*/, advCoef_3rd /*E*16*/;
  bool fct_ready = false;
};

// topology (orc_mesh.cc)
void read_connectivity_2D(const int32_t *buf, size_t dim1, size_t dim2, Crs &out);
void init_cnx(const Crs &file, Crs &out);
void transpose_cnx(const Crs &c2v, int n_to, Crs &v2c);
void finalize_mesh(Mesh &m);

// SW tasks (orc_sw.cc)
void sw_compute_solve_diagnostics(const Mesh &m, int nL, const double *h,
                                  const double *u, double *h_edge,
                                  double *h_vertex, double *circulation,
                                  double *vorticity, double *ke,
                                  double *pv_edge, double *pv_vertex);
void sw_init_provis(const Mesh &m, int nL, int nT, const double *u_old,
                    const double *h_old, const double *tr_old,
                    const double *h_edge_old, const double *ke_old,
                    const double *pv_edge_old, double *u_p, double *h_p,
                    double *tr_p, double *h_edge_p, double *ke_p,
                    double *pv_edge_p);
void sw_setup_timestep(const Mesh &m, int nL, int nT, const double *u_old,
                       const double *h_old, const double *tr_old, double *u_new,
                       double *h_new, double *tr_new);
void sw_compute_tend(const Mesh &m, int nL, double gravity,
                     const double *h_edge, const double *h, const double *h_s,
                     const double *u, const double *ke, const double *pv_edge,
                     double *tend_h, double *tend_u);
void sw_compute_scalar_tend(const Mesh &m, int nL, int nT,
                            const double *h_edge, const double *u,
                            const double *tracers, double *tracer_tend);
void sw_compute_substep(const Mesh &m, int nL, int nT, int freeze_u,
                        double *u_p, const double *u_old, const double *u_tend,
                        double *h_p, const double *h_old, const double *h_tend,
                        double *tr_p, const double *tr_old,
                        const double *tr_tend, double rkweight);
void sw_accumulate_update(const Mesh &m, int nL, int nT, double *u_new,
                          const double *u_tend, double *h_new,
                          const double *h_tend, double *tr_new,
                          const double *tr_tend, double rkweight);
void sw_finalize_timestep(const Mesh &m, int nL, int nT, int freeze_u,
                          const double *u_old, double *u_new,
                          const double *h_new, double *tr_new);
void sw_timeshift(const Mesh &m, int nL, int nT, double *u_old, double *h_old,
                  double *tr_old, double *h_edge_old, double *ke_old,
                  double *pv_edge_old, const double *u_new, const double *h_new,
                  const double *tr_new, const double *h_edge_new,
                  const double *ke_new, const double *pv_edge_new);
double sw_compute_error(const Mesh &m, int nL, double a, const double *h,
                        double elapsed);
// del2 viscosity on u: build-defined extension (no reference counterpart)
void sw_del2_viscosity(const Mesh &m, int nL, double visc2, const double *u,
                       const double *vorticity, double *divergence,
                       double *tend_u);

// ocean tasks (orc_ocean.cc)
void ocn_copy_tracers(const Mesh &m, int nL, int nT, const double *tr,
                      double *trNew, double *trPrev);
void ocn_reinitialize_tracer_tendencies(const Mesh &m, int nL, int nT,
                                        double *tend);
void ocn_calc_normal_thickness_flux(const Mesh &m, int nL,
                                    const double *layerThicknessEdge,
                                    const double *normalVelocity,
                                    double *normalThicknessFlux);
void ocn_calc_provisional_layer_thicknesses(const Mesh &m, int nL,
                                            const double *layerThickness,
                                            const double *normalThicknessFlux,
                                            double *hProv, double *hProvInv,
                                            double *hNewInv, double dt);
void ocn_low_order(const Mesh &m, int nL, int nT,
                   const double *normalThicknessFlux, const double *tracers,
                   double *lowOrderFlux);
void ocn_high_order(const Mesh &m, int nL, int nT, double coef3rd,
                    const double *normalThicknessFlux, const double *tracers,
                    double *highOrderFlux);
void ocn_correct_and_rescale(const Mesh &m, int nL, int nT,
                             const double *hProvInv,
                             const double *layerThickness,
                             const double *normalThicknessFlux,
                             const double *lowOrderFlux, const double *tracers,
                             double *highOrderFlux, double *flxIn,
                             double *flxOut, double *tend, double *tracerMax,
                             double *tracerMin, double dt);
void ocn_accumulate_fct(const Mesh &m, int nL, int nT, const double *hProvInv,
                        const double *layerThickness,
                        const double *highOrderFlux, const double *tracers,
                        double *tracersProv, double *tend, double dt);
void ocn_hmix_del2(const Mesh &m, int nL, int nT,
                   const double *layerThicknessEdge, const double *tracers,
                   double *tend, double coef);
void ocn_step_tracers_forward(const Mesh &m, int nL, int nT, const double *tend,
                              const double *tracers, double *tracersNew,
                              double *tracersPrev, double rk_weight,
                              double rk_subweight);
void ocn_update_tracers(const Mesh &m, int nL, int nT, const double *trNew,
                        double *tr);

// FCT setup (orc_setup.cc)
void setup_fct(Mesh &m, int nL, double a);

// The inverse of the n x n row-major matrix a by migs (matrix.hh:137-179)
// into x: exposed so the tests can check it against a library inverse.
void migs_dense(int n, const double *a, double *x);

// drivers (orc_driver.cc)
struct SwState;  // opaque
}  // namespace orc
