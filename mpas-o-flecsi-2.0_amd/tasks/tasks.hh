// C++ task layer over the C ABI: the reference's FleCSI tasks for the
// shallow-water / FCT path (apps/sw/tasks/solve.hh:12-136,
// apps/ocean/tasks/solve.hh:12-111) with the same names and argument order.
//
//   reference                                  here
//   mesh::accessor<ro, ro> m                   mpas::hip::mesh &m  (owns the
//                                              flattened topology and the
//                                              static mesh fields on device)
//   field<T>::accessor<P, P> x                 mpas::hip::field &x (device array
//                                              in the reference layout)
//   static mesh fields (dvEdge, areaCell, ...) held by the mesh, not passed
//   flog_fatal / flog_assert (abort)           throws mpas::hip::error
//
// Calls are asynchronous on the mesh's stream, as FleCSI tasks are ordered
// by the runtime; results are visible to the host after m.sync().
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpas_hip.h"

namespace mpas { namespace hip {

struct error : std::runtime_error {
  int code;
  error(int c, const std::string &what) : std::runtime_error(what), code(c) {}
};

// Throws on a non-zero status of the C ABI.
void check(int rc, const char *where);

// The mesh topology slot (mesh::initialize + the mesh fields).
class mesh {
 public:
  mesh(const mpas_mesh_desc &d, const mpas_params &p);
  ~mesh();
  mesh(const mesh &) = delete;
  mesh &operator=(const mesh &) = delete;
  mpas_ctx *ctx() const { return ctx_; }
  int64_t cells() const { return dims_[0]; }
  int64_t edges() const { return dims_[1]; }
  int64_t vertices() const { return dims_[2]; }
  int64_t levels() const { return dims_[3]; }
  int64_t tracers() const { return dims_[4]; }
  const mpas_params &params() const { return params_; }
  void sync() const;

 private:
  mpas_ctx *ctx_ = nullptr;
  mpas_params params_;
  int64_t dims_[5] = {};
};

// A device field (the storage FleCSI owns in the reference).
enum class entity { cells, edges, vertices };
class field {
 public:
  field() = default;
  // n_per = values per entity (nLevels, nLevels * nTracers, or 1)
  field(mesh &m, entity where, int64_t n_per);
  ~field();
  field(field &&o) noexcept;
  field &operator=(field &&o) noexcept;
  field(const field &) = delete;
  field &operator=(const field &) = delete;
  double *data() const { return d_; }
  int64_t size() const { return n_; }
  void upload(const std::vector<double> &host);
  std::vector<double> download() const;

 private:
  mesh *m_ = nullptr;
  double *d_ = nullptr;
  int64_t n_ = 0;
};

namespace sw { namespace task {
// apps/sw/tasks/solve.hh
void compute_solve_diagnostics(mesh &m, const field &h, const field &u, field &h_edge,
                               field &h_vertex, field &circulation, field &vorticity,
                               field &ke, field &pv_edge, field &pv_vertex);
void init_provis(mesh &m, const field &u_old, const field &h_old, const field &tracers_old,
                 const field &h_edge_old, const field &ke_old, const field &pv_edge_old,
                 field &u_provis, field &h_provis, field &tracers_provis, field &h_edge_provis,
                 field &ke_provis, field &pv_edge_provis);
void setup_timestep(mesh &m, const field &u_old, const field &h_old, const field &tracers_old,
                    field &u_new, field &h_new, field &tracers_new);
void compute_tend(mesh &m, const field &h_edge, const field &h, const field &h_s,
                  const field &u, const field &ke, const field &pv_edge, field &tend_h,
                  field &tend_u);
void compute_scalar_tend(mesh &m, const field &h_edge, const field &u, const field &tracers,
                         field &tracer_tend);
void compute_substep(mesh &m, field &u_provis, const field &u_old, const field &u_tend,
                     field &h_provis, const field &h_old, const field &h_tend,
                     field &tracers_provis, const field &tracers_old,
                     const field &tracers_tend, double rkweight);
void accumulate_update(mesh &m, field &u_new, const field &u_tend, field &h_new,
                       const field &h_tend, field &tracers_new, const field &tracers_tend,
                       double rkweight);
void finalize_timestep(mesh &m, const field &u_old, field &u_new, const field &h_new,
                       field &tracers_new);
void timeshift(mesh &m, field &u_old, field &h_old, field &tracers_old, field &h_edge_old,
               field &ke_old, field &pv_edge_old, const field &u_new, const field &h_new,
               const field &tracers_new, const field &h_edge_new, const field &ke_new,
               const field &pv_edge_new);
double compute_error(mesh &m, const field &h, const field &latCell, const field &lonCell,
                     double elapsed);
}}  // namespace sw::task

namespace ocean { namespace task {
// apps/ocean/tasks/solve.hh
void copy_tracers(mesh &m, const field &tracers, field &tracersNew, field &tracersPrev);
void reinitialize_tracer_tendencies(mesh &m, field &tend);
void calc_normal_thickness_flux(mesh &m, const field &layerThicknessEdge,
                                const field &normalVelocity, field &normalThicknessFlux);
void calc_provisional_layer_thicknesses(mesh &m, const field &layerThickness,
                                        const field &normalThicknessFlux, field &hProv,
                                        field &hProvInv, field &hNewInv, double dt);
void calc_horiz_tracer_advection_flux_mono_low_order(mesh &m, const field &normalThicknessFlux,
                                                     const field &tracers, field &lowOrderFlux);
void calc_horiz_tracer_advection_flux_mono_high_order(mesh &m,
                                                      const field &normalThicknessFlux,
                                                      const field &tracers,
                                                      field &highOrderFlux);
void correct_and_rescale_high_order_mono_horiz_flux(
    mesh &m, const field &hProvInv, const field &layerThickness,
    const field &normalThicknessFlux, const field &lowOrderFlux, const field &tracers,
    field &highOrderFlux, field &flxIn, field &flxOut, field &tend, field &tracerMax,
    field &tracerMin, double dt);
void accumulate_horiz_tracer_advection_tend_mono_FCT(mesh &m, const field &hProvInv,
                                                     const field &layerThickness,
                                                     const field &highOrderFlux,
                                                     const field &tracers,
                                                     field &tracersProv, field &tend,
                                                     double dt);
void tracer_hmix_del2_tend(mesh &m, const field &layerThicknessEdge, const field &tracers,
                           field &tend, double coef);
void step_tracers_forward(mesh &m, const field &tend, const field &tracers, field &tracersNew,
                          field &tracersPrev, double rk_weight, double rk_subweight);
void update_tracers(mesh &m, const field &tracersNew, field &tracers);
}}  // namespace ocean::task

}}  // namespace mpas::hip
