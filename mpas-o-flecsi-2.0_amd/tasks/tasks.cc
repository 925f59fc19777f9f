// C++ task layer over the C ABI (see tasks.hh).
#include "tasks.hh"

namespace mpas { namespace hip {

void check(int rc, const char *where) {
  if (rc != 0) throw error(rc, std::string(where) + ": " + mpas_last_error());
}

mesh::mesh(const mpas_mesh_desc &d, const mpas_params &p) : params_(p) {
  check(mpas_ctx_create(&d, &p, &ctx_), "mpas_ctx_create");
  check(mpas_ctx_dims(ctx_, dims_), "mpas_ctx_dims");
}

mesh::~mesh() { mpas_ctx_destroy(ctx_); }

void mesh::sync() const { check(mpas_ctx_sync(ctx_), "mpas_ctx_sync"); }

field::field(mesh &m, entity where, int64_t n_per) : m_(&m) {
  const int64_t n = where == entity::cells ? m.cells()
                  : where == entity::edges ? m.edges() : m.vertices();
  n_ = n * n_per;
  void *p = nullptr;
  if (n_) check(mpas_buf_alloc(m.ctx(), sizeof(double) * n_, &p), "mpas_buf_alloc");
  d_ = static_cast<double *>(p);
}

field::~field() {
  if (d_ && m_) mpas_buf_free(m_->ctx(), d_);
}

field::field(field &&o) noexcept : m_(o.m_), d_(o.d_), n_(o.n_) {
  o.d_ = nullptr;
  o.n_ = 0;
}

field &field::operator=(field &&o) noexcept {
  if (this != &o) {
    if (d_ && m_) mpas_buf_free(m_->ctx(), d_);
    m_ = o.m_;
    d_ = o.d_;
    n_ = o.n_;
    o.d_ = nullptr;
    o.n_ = 0;
  }
  return *this;
}

void field::upload(const std::vector<double> &host) {
  if (int64_t(host.size()) != n_) throw error(MPAS_ESHAPE, "field::upload: size mismatch");
  if (n_) check(mpas_buf_upload(m_->ctx(), d_, host.data(), sizeof(double) * n_), "upload");
}

std::vector<double> field::download() const {
  std::vector<double> out(n_);
  if (n_) check(mpas_buf_download(m_->ctx(), out.data(), d_, sizeof(double) * n_), "download");
  return out;
}

namespace sw { namespace task {

void compute_solve_diagnostics(mesh &m, const field &h, const field &u, field &h_edge,
                               field &h_vertex, field &circulation, field &vorticity,
                               field &ke, field &pv_edge, field &pv_vertex) {
  check(mpas_sw_compute_solve_diagnostics(m.ctx(), h.data(), u.data(), h_edge.data(),
                                          h_vertex.data(), circulation.data(),
                                          vorticity.data(), ke.data(), pv_edge.data(),
                                          pv_vertex.data()),
        "compute_solve_diagnostics");
}

void init_provis(mesh &m, const field &u_old, const field &h_old, const field &tracers_old,
                 const field &h_edge_old, const field &ke_old, const field &pv_edge_old,
                 field &u_provis, field &h_provis, field &tracers_provis, field &h_edge_provis,
                 field &ke_provis, field &pv_edge_provis) {
  check(mpas_sw_init_provis(m.ctx(), u_old.data(), h_old.data(), tracers_old.data(),
                            h_edge_old.data(), ke_old.data(), pv_edge_old.data(),
                            u_provis.data(), h_provis.data(), tracers_provis.data(),
                            h_edge_provis.data(), ke_provis.data(), pv_edge_provis.data()),
        "init_provis");
}

void setup_timestep(mesh &m, const field &u_old, const field &h_old, const field &tracers_old,
                    field &u_new, field &h_new, field &tracers_new) {
  check(mpas_sw_setup_timestep(m.ctx(), u_old.data(), h_old.data(), tracers_old.data(),
                               u_new.data(), h_new.data(), tracers_new.data()),
        "setup_timestep");
}

void compute_tend(mesh &m, const field &h_edge, const field &h, const field &h_s,
                  const field &u, const field &ke, const field &pv_edge, field &tend_h,
                  field &tend_u) {
  check(mpas_sw_compute_tend(m.ctx(), h_edge.data(), h.data(), h_s.data(), u.data(), ke.data(),
                             pv_edge.data(), tend_h.data(), tend_u.data()),
        "compute_tend");
}

void compute_scalar_tend(mesh &m, const field &h_edge, const field &u, const field &tracers,
                         field &tracer_tend) {
  check(mpas_sw_compute_scalar_tend(m.ctx(), h_edge.data(), u.data(), tracers.data(),
                                    tracer_tend.data()),
        "compute_scalar_tend");
}

void compute_substep(mesh &m, field &u_provis, const field &u_old, const field &u_tend,
                     field &h_provis, const field &h_old, const field &h_tend,
                     field &tracers_provis, const field &tracers_old,
                     const field &tracers_tend, double rkweight) {
  check(mpas_sw_compute_substep(m.ctx(), u_provis.data(), u_old.data(), u_tend.data(),
                                h_provis.data(), h_old.data(), h_tend.data(),
                                tracers_provis.data(), tracers_old.data(), tracers_tend.data(),
                                rkweight),
        "compute_substep");
}

void accumulate_update(mesh &m, field &u_new, const field &u_tend, field &h_new,
                       const field &h_tend, field &tracers_new, const field &tracers_tend,
                       double rkweight) {
  check(mpas_sw_accumulate_update(m.ctx(), u_new.data(), u_tend.data(), h_new.data(),
                                  h_tend.data(), tracers_new.data(), tracers_tend.data(),
                                  rkweight),
        "accumulate_update");
}

void finalize_timestep(mesh &m, const field &u_old, field &u_new, const field &h_new,
                       field &tracers_new) {
  check(mpas_sw_finalize_timestep(m.ctx(), u_old.data(), u_new.data(), h_new.data(),
                                  tracers_new.data()),
        "finalize_timestep");
}

void timeshift(mesh &m, field &u_old, field &h_old, field &tracers_old, field &h_edge_old,
               field &ke_old, field &pv_edge_old, const field &u_new, const field &h_new,
               const field &tracers_new, const field &h_edge_new, const field &ke_new,
               const field &pv_edge_new) {
  check(mpas_sw_timeshift(m.ctx(), u_old.data(), h_old.data(), tracers_old.data(),
                          h_edge_old.data(), ke_old.data(), pv_edge_old.data(), u_new.data(),
                          h_new.data(), tracers_new.data(), h_edge_new.data(), ke_new.data(),
                          pv_edge_new.data()),
        "timeshift");
}

double compute_error(mesh &m, const field &h, const field &latCell, const field &lonCell,
                     double elapsed) {
  double err = 0;
  check(mpas_sw_compute_error(m.ctx(), h.data(), latCell.data(), lonCell.data(),
                              m.params().sphere_radius, elapsed, &err),
        "compute_error");
  return err;
}

}}  // namespace sw::task

namespace ocean { namespace task {

void copy_tracers(mesh &m, const field &tracers, field &tracersNew, field &tracersPrev) {
  check(mpas_fct_copy_tracers(m.ctx(), tracers.data(), tracersNew.data(), tracersPrev.data()),
        "copy_tracers");
}

void reinitialize_tracer_tendencies(mesh &m, field &tend) {
  check(mpas_fct_reinitialize_tracer_tendencies(m.ctx(), tend.data()),
        "reinitialize_tracer_tendencies");
}

void calc_normal_thickness_flux(mesh &m, const field &layerThicknessEdge,
                                const field &normalVelocity, field &normalThicknessFlux) {
  check(mpas_fct_calc_normal_thickness_flux(m.ctx(), layerThicknessEdge.data(),
                                            normalVelocity.data(), normalThicknessFlux.data()),
        "calc_normal_thickness_flux");
}

void calc_provisional_layer_thicknesses(mesh &m, const field &layerThickness,
                                        const field &normalThicknessFlux, field &hProv,
                                        field &hProvInv, field &hNewInv, double dt) {
  check(mpas_fct_calc_provisional_layer_thicknesses(m.ctx(), layerThickness.data(),
                                                    normalThicknessFlux.data(), hProv.data(),
                                                    hProvInv.data(), hNewInv.data(), dt),
        "calc_provisional_layer_thicknesses");
}

void calc_horiz_tracer_advection_flux_mono_low_order(mesh &m, const field &normalThicknessFlux,
                                                     const field &tracers, field &lowOrderFlux) {
  check(mpas_fct_low_order(m.ctx(), normalThicknessFlux.data(), tracers.data(),
                           lowOrderFlux.data()),
        "calc_horiz_tracer_advection_flux_mono_low_order");
}

void calc_horiz_tracer_advection_flux_mono_high_order(mesh &m,
                                                      const field &normalThicknessFlux,
                                                      const field &tracers,
                                                      field &highOrderFlux) {
  check(mpas_fct_high_order(m.ctx(), normalThicknessFlux.data(), tracers.data(),
                            highOrderFlux.data()),
        "calc_horiz_tracer_advection_flux_mono_high_order");
}

void correct_and_rescale_high_order_mono_horiz_flux(
    mesh &m, const field &hProvInv, const field &layerThickness,
    const field &normalThicknessFlux, const field &lowOrderFlux, const field &tracers,
    field &highOrderFlux, field &flxIn, field &flxOut, field &tend, field &tracerMax,
    field &tracerMin, double dt) {
  check(mpas_fct_correct_and_rescale(m.ctx(), hProvInv.data(), layerThickness.data(),
                                     normalThicknessFlux.data(), lowOrderFlux.data(),
                                     tracers.data(), highOrderFlux.data(), flxIn.data(),
                                     flxOut.data(), tend.data(), tracerMax.data(),
                                     tracerMin.data(), dt),
        "correct_and_rescale_high_order_mono_horiz_flux");
}

void accumulate_horiz_tracer_advection_tend_mono_FCT(mesh &m, const field &hProvInv,
                                                     const field &layerThickness,
                                                     const field &highOrderFlux,
                                                     const field &tracers,
                                                     field &tracersProv, field &tend,
                                                     double dt) {
  check(mpas_fct_accumulate_tend(m.ctx(), hProvInv.data(), layerThickness.data(),
                                 highOrderFlux.data(), tracers.data(), tracersProv.data(),
                                 tend.data(), dt),
        "accumulate_horiz_tracer_advection_tend_mono_FCT");
}

void tracer_hmix_del2_tend(mesh &m, const field &layerThicknessEdge, const field &tracers,
                           field &tend, double coef) {
  check(mpas_hmix_del2_tend(m.ctx(), layerThicknessEdge.data(), tracers.data(), tend.data(),
                            coef),
        "tracer_hmix_del2_tend");
}

void step_tracers_forward(mesh &m, const field &tend, const field &tracers, field &tracersNew,
                          field &tracersPrev, double rk_weight, double rk_subweight) {
  check(mpas_fct_step_tracers_forward(m.ctx(), tend.data(), tracers.data(), tracersNew.data(),
                                      tracersPrev.data(), rk_weight, rk_subweight),
        "step_tracers_forward");
}

void update_tracers(mesh &m, const field &tracersNew, field &tracers) {
  check(mpas_fct_update_tracers(m.ctx(), tracersNew.data(), tracers.data()), "update_tracers");
}

}}  // namespace ocean::task

}}  // namespace mpas::hip
