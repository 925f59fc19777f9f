// ORACLE — test infrastructure only (see orc_internal.h).
// Drivers restated from apps/sw/run.cc:13-113 and apps/ocean/run.cc:10-77,
// plus a flat C API for the Python tests and the CPU-baseline timing.
#include <chrono>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>

#include "orc_internal.h"

#ifdef _OPENMP
#include <omp.h>
#endif

using namespace orc;

namespace {
thread_local std::string g_err;

struct OrcParams {  // mirrors orc_params in oracle.h
  int32_t nL, nT, tracer_adv, freeze_u;
  double gravity, a, coef3rd, visc2, kappa2;
};

Params to_params(const OrcParams *q) {
  Params p;
  p.nL = q->nL;
  p.nT = q->nT;
  p.tracer_adv = q->tracer_adv;
  p.freeze_u = q->freeze_u;
  p.gravity = q->gravity;
  p.a = q->a;
  p.coef3rd = q->coef3rd;
  p.visc2 = q->visc2;
  p.kappa2 = q->kappa2;
  return p;
}

// Shallow-water state with the reference's time versions
// (apps/sw/state.hh:74: prev = curr = 0, next = 1, provis = 2).
struct SwRun {
  Mesh *m;
  Params p;
  enum { PREV = 0, NEXT = 1, PROVIS = 2 };
  std::vector<double> h[3], u[3], h_edge[3], h_vertex[3], ke[3], pv_edge[3], tr[3];
  std::vector<double> circulation, vorticity, pv_vertex, tend_h, tend_u, tend_tr, h_s, divergence;
  std::vector<double> ntf, hProv, hProvInv, hNewInv, lo, ho, flxIn, flxOut, trMax, trMin, trProv;
  double elapsed = 0;

  SwRun(Mesh *mesh, const Params &pp) : m(mesh), p(pp) {
    size_t C = m->nC, E = m->nE, V = m->nV, L = p.nL, T = p.nT;
    for (int v = 0; v < 3; v++) {
      h[v].assign(C * L, 0);
      u[v].assign(E * L, 0);
      h_edge[v].assign(E * L, 0);
      h_vertex[v].assign(V * L, 0);
      ke[v].assign(C * L, 0);
      pv_edge[v].assign(E * L, 0);
      tr[v].assign(C * L * T, 0);
    }
    circulation.assign(V * L, 0);
    vorticity.assign(V * L, 0);
    pv_vertex.assign(V * L, 0);
    tend_h.assign(C * L, 0);
    tend_u.assign(E * L, 0);
    tend_tr.assign(C * L * T, 0);
    h_s.assign(C, 0);
    divergence.assign(C * L, 0);
    ntf.assign(E * L, 0);
    hProv.assign(C * L, 0);
    hProvInv.assign(C * L, 0);
    hNewInv.assign(C * L, 0);
    lo.assign(E * L * T, 0);
    ho.assign(E * L * T, 0);
    flxIn.assign(C * L * T, 0);
    flxOut.assign(C * L * T, 0);
    trMax.assign(C * L * T, 0);
    trMin.assign(C * L * T, 0);
    trProv.assign(C * L * T, 0);
  }

  void diagnostics(int v) {
    sw_compute_solve_diagnostics(*m, p.nL, h[v].data(), u[v].data(), h_edge[v].data(),
                                 h_vertex[v].data(), circulation.data(), vorticity.data(),
                                 ke[v].data(), pv_edge[v].data(), pv_vertex.data());
  }

  // Tracer tendency of one stage.  The reference SW driver uses the centred
  // compute_scalar_tend (apps/sw/run.cc:33-35).  With tracer_adv == 2 the
  // ocean FCT tasks (apps/ocean/run.cc:19-42) are driven from the provisional
  // SW state instead: layerThicknessEdge = h_edge[provis], normalVelocity =
  // u[provis], layerThickness = h[provis], tracers[prev] = tracers[provis].
  void tracer_tend(double dt) {
    const int nL = p.nL, nT = p.nT;
    if (nT == 0) return;
    const int P = PROVIS;
    if (p.tracer_adv == 2) {
      ocn_reinitialize_tracer_tendencies(*m, nL, nT, tend_tr.data());
      ocn_calc_normal_thickness_flux(*m, nL, h_edge[P].data(), u[P].data(), ntf.data());
      ocn_calc_provisional_layer_thicknesses(*m, nL, h[P].data(), ntf.data(), hProv.data(),
                                             hProvInv.data(), hNewInv.data(), dt);
      ocn_low_order(*m, nL, nT, ntf.data(), tr[P].data(), lo.data());
      ocn_high_order(*m, nL, nT, p.coef3rd, ntf.data(), tr[P].data(), ho.data());
      ocn_correct_and_rescale(*m, nL, nT, hProvInv.data(), h[P].data(), ntf.data(), lo.data(),
                              tr[P].data(), ho.data(), flxIn.data(), flxOut.data(),
                              tend_tr.data(), trMax.data(), trMin.data(), dt);
      ocn_accumulate_fct(*m, nL, nT, hProvInv.data(), h[P].data(), ho.data(), tr[P].data(),
                         trProv.data(), tend_tr.data(), dt);
    } else {
      sw_compute_scalar_tend(*m, nL, nT, h_edge[P].data(), u[P].data(), tr[P].data(),
                             tend_tr.data());
    }
    if (p.kappa2 != 0.0)
      ocn_hmix_del2(*m, nL, nT, h_edge[P].data(), tr[P].data(), tend_tr.data(), p.kappa2);
  }

  // apps/sw/run.cc:13-70
  void advance_timestep(double dt) {
    const int nL = p.nL, nT = p.nT;
    const int A = PREV, N = NEXT, P = PROVIS;
    sw_init_provis(*m, nL, nT, u[A].data(), h[A].data(), tr[A].data(), h_edge[A].data(),
                   ke[A].data(), pv_edge[A].data(), u[P].data(), h[P].data(), tr[P].data(),
                   h_edge[P].data(), ke[P].data(), pv_edge[P].data());
    sw_setup_timestep(*m, nL, nT, u[A].data(), h[A].data(), tr[A].data(), u[N].data(),
                      h[N].data(), tr[N].data());
    const double rk_weights[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
    const double rk_substep_weights[4] = {dt / 2, dt / 2, dt, 0};
    for (int rkstep = 0; rkstep < 4; rkstep++) {
      sw_compute_tend(*m, nL, p.gravity, h_edge[P].data(), h[P].data(), h_s.data(), u[P].data(),
                      ke[P].data(), pv_edge[P].data(), tend_h.data(), tend_u.data());
      if (p.visc2 != 0.0)
        sw_del2_viscosity(*m, nL, p.visc2, u[P].data(), vorticity.data(), divergence.data(),
                          tend_u.data());
      tracer_tend(dt);
      if (rkstep < 3) {
        sw_compute_substep(*m, nL, nT, p.freeze_u, u[P].data(), u[A].data(), tend_u.data(),
                           h[P].data(), h[A].data(), tend_h.data(), tr[P].data(),
                           tr[A].data(), tend_tr.data(), rk_substep_weights[rkstep]);
        diagnostics(P);
      }
      sw_accumulate_update(*m, nL, nT, u[N].data(), tend_u.data(), h[N].data(),
                           tend_h.data(), tr[N].data(), tend_tr.data(), rk_weights[rkstep]);
    }
    sw_finalize_timestep(*m, nL, nT, p.freeze_u, u[A].data(), u[N].data(), h[N].data(),
                         tr[N].data());
    diagnostics(N);
    sw_timeshift(*m, nL, nT, u[A].data(), h[A].data(), tr[A].data(), h_edge[A].data(),
                 ke[A].data(), pv_edge[A].data(), u[N].data(), h[N].data(), tr[N].data(),
                 h_edge[N].data(), ke[N].data(), pv_edge[N].data());
  }
};

// Tracer-only ocean run (apps/ocean/state.hh:146: prev, curr, next, prov).
struct OcnRun {
  Mesh *m;
  Params p;
  int mono = 1, hmix = 0;
  double hmix_coef = 0;
  std::vector<double> tr[4], layerThickness[4], normalVelocity, layerThicknessEdge, tend;
  std::vector<double> ntf, hProvInv, hNewInv, lo, ho, flxIn, flxOut, trMax, trMin;
  OcnRun(Mesh *mesh, const Params &pp) : m(mesh), p(pp) {
    size_t C = m->nC, E = m->nE, L = p.nL, T = p.nT;
    for (int v = 0; v < 4; v++) {
      tr[v].assign(C * L * T, 0);
      layerThickness[v].assign(C * L, 0);
    }
    normalVelocity.assign(E * L, 0);
    layerThicknessEdge.assign(E * L, 0);
    tend.assign(C * L * T, 0);
    ntf.assign(E * L, 0);
    hProvInv.assign(C * L, 0);
    hNewInv.assign(C * L, 0);
    lo.assign(E * L * T, 0);
    ho.assign(E * L * T, 0);
    flxIn.assign(C * L * T, 0);
    flxOut.assign(C * L * T, 0);
    trMax.assign(C * L * T, 0);
    trMin.assign(C * L * T, 0);
  }
  // apps/ocean/initialize.cc:278-279 -> init_functions.cc:30-45
  void init_layerThicknessEdge() {
    const int nL = p.nL;
    const double *lt = layerThickness[1].data();
    for (int e = 0; e < m->nE; e++) {
      for (int k = 0; k < nL; k++) layerThicknessEdge[e * nL + k] = 0.0;
      int c1 = m->e2c.at(e, 0), c2 = m->e2c.at(e, 1);
      for (int k = 0; k < nL; k++)
        layerThicknessEdge[e * nL + k] = 0.5 * (lt[c1 * nL + k] + lt[c2 * nL + k]);
    }
  }
  // apps/ocean/run.cc:10-50
  void rk4_timestep(double dt) {
    const int nL = p.nL, nT = p.nT;
    enum { PREV = 0, CURR = 1, NEXT = 2, PROV = 3 };
    ocn_copy_tracers(*m, nL, nT, tr[CURR].data(), tr[NEXT].data(), tr[PREV].data());
    const double rk_weights[4] = {dt / 6, dt / 3, dt / 3, dt / 6};
    const double rk_substep_weights[4] = {dt / 2, dt / 2, dt, 0};
    for (int rk = 0; rk < 4; rk++) {
      ocn_reinitialize_tracer_tendencies(*m, nL, nT, tend.data());
      if (mono) {
        ocn_calc_normal_thickness_flux(*m, nL, layerThicknessEdge.data(), normalVelocity.data(),
                                       ntf.data());
        ocn_calc_provisional_layer_thicknesses(*m, nL, layerThickness[CURR].data(), ntf.data(),
                                               layerThickness[PROV].data(), hProvInv.data(),
                                               hNewInv.data(), dt);
        ocn_low_order(*m, nL, nT, ntf.data(), tr[PREV].data(), lo.data());
        ocn_high_order(*m, nL, nT, p.coef3rd, ntf.data(), tr[PREV].data(), ho.data());
        ocn_correct_and_rescale(*m, nL, nT, hProvInv.data(), layerThickness[CURR].data(),
                                ntf.data(), lo.data(), tr[PREV].data(), ho.data(),
                                flxIn.data(), flxOut.data(), tend.data(), trMax.data(),
                                trMin.data(), dt);
        ocn_accumulate_fct(*m, nL, nT, hProvInv.data(), layerThickness[CURR].data(), ho.data(),
                           tr[PREV].data(), tr[PROV].data(), tend.data(), dt);
      }
      if (hmix)
        ocn_hmix_del2(*m, nL, nT, layerThicknessEdge.data(), tr[PREV].data(), tend.data(),
                      hmix_coef);
      ocn_step_tracers_forward(*m, nL, nT, tend.data(), tr[CURR].data(), tr[NEXT].data(),
                               tr[PREV].data(), rk_weights[rk], rk_substep_weights[rk]);
    }
    ocn_update_tracers(*m, nL, nT, tr[NEXT].data(), tr[CURR].data());
  }
};

template <class F>
int guard(F &&f) {
  try {
    f();
    return 0;
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return -1;
  }
}

std::vector<double> *sw_field(SwRun *s, const std::string &name) {
  static const std::map<std::string, int> ver = {{"prev", 0}, {"next", 1}, {"provis", 2}};
  if (name == "circulation") return &s->circulation;
  if (name == "vorticity") return &s->vorticity;
  if (name == "pv_vertex") return &s->pv_vertex;
  if (name == "tend_h") return &s->tend_h;
  if (name == "tend_u") return &s->tend_u;
  if (name == "tend_tracers") return &s->tend_tr;
  if (name == "h_s") return &s->h_s;
  if (name == "divergence") return &s->divergence;
  auto dot = name.find('.');
  std::string base = name.substr(0, dot);
  int v = 0;
  if (dot != std::string::npos) {
    auto it = ver.find(name.substr(dot + 1));
    if (it == ver.end()) return nullptr;
    v = it->second;
  }
  if (base == "h") return &s->h[v];
  if (base == "u") return &s->u[v];
  if (base == "h_edge") return &s->h_edge[v];
  if (base == "h_vertex") return &s->h_vertex[v];
  if (base == "ke") return &s->ke[v];
  if (base == "pv_edge") return &s->pv_edge[v];
  if (base == "tracers") return &s->tr[v];
  return nullptr;
}

}  // namespace

extern "C" {

const char *orc_last_error(void) { return g_err.c_str(); }

void *orc_mesh_new(int nC, int nE, int nV, int maxEdges, int maxEdges2) {
  auto *m = new Mesh();
  m->nC = nC;
  m->nE = nE;
  m->nV = nV;
  m->maxEdges = maxEdges;
  m->maxEdges2 = maxEdges2;
  return m;
}

void orc_mesh_free(void *h) { delete static_cast<Mesh *>(h); }

int orc_mesh_set_int(void *h, const char *name, const int32_t *p, int64_t n) {
  return guard([&] {
    Mesh &m = *static_cast<Mesh *>(h);
    std::string s(name);
    std::vector<int32_t> v(p, p + n);
    if (s == "verticesOnCell") m.f_voc = v;
    else if (s == "edgesOnCell") m.f_eoc = v;
    else if (s == "cellsOnCell") m.f_coc = v;
    else if (s == "verticesOnEdge") m.f_voe = v;
    else if (s == "edgesOnEdge") m.f_eoe = v;
    else if (s == "maxLevelCell") m.maxLevelCell.assign(p, p + n);
    else if (s == "boundaryCell") m.boundaryCell.assign(p, p + n);
    else throw std::runtime_error("unknown int field " + s);
  });
}

int orc_mesh_set_dbl(void *h, const char *name, const double *p, int64_t n) {
  return guard([&] {
    Mesh &m = *static_cast<Mesh *>(h);
    std::string s(name);
    std::vector<double> v(p, p + n);
    if (s == "areaCell") m.areaCell = v;
    else if (s == "dvEdge") m.dvEdge = v;
    else if (s == "dcEdge") m.dcEdge = v;
    else if (s == "areaTriangle") m.areaTriangle = v;
    else if (s == "kiteAreasOnVertex") m.kite = v;
    else if (s == "fVertex") m.fVertex = v;
    else if (s == "weightsOnEdge") {
      // mpasoflecsi/io/helpers.hh:172-184 reads 16 columns at the file stride
      if (m.maxEdges2 != 16) throw std::runtime_error("weightsOnEdge stride must be 16");
      m.weightsOnEdge = v;
    } else if (s == "xCell") m.xCell = v;
    else if (s == "yCell") m.yCell = v;
    else if (s == "zCell") m.zCell = v;
    else if (s == "xVertex") m.xVertex = v;
    else if (s == "yVertex") m.yVertex = v;
    else if (s == "zVertex") m.zVertex = v;
    else if (s == "latCell") m.latCell = v;
    else if (s == "lonCell") m.lonCell = v;
    else if (s == "latEdge") m.latEdge = v;
    else if (s == "lonEdge") m.lonEdge = v;
    else if (s == "latVertex") m.latVertex = v;
    else if (s == "lonVertex") m.lonVertex = v;
    else if (s == "angleEdge") m.angleEdge = v;
    else if (s == "meshDensity") m.meshDensity = v;
    else throw std::runtime_error("unknown double field " + s);
  });
}

int orc_mesh_finalize(void *h) {
  return guard([&] { finalize_mesh(*static_cast<Mesh *>(h)); });
}

// Threads of the gather loops (ORC_PAR); returns the count in use.
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

int orc_migs(int n, const double *a, double *x) {
  return guard([&] { migs_dense(n, a, x); });
}

int orc_mesh_setup_fct(void *h, int nL, double a) {
  return guard([&] { setup_fct(*static_cast<Mesh *>(h), nL, a); });
}

// connectivity lists: c2v c2e c2c e2v e2e v2c e2c v2e
int orc_mesh_cnx(void *h, const char *which, const int **off, const int **idx, int *rows) {
  return guard([&] {
    Mesh &m = *static_cast<Mesh *>(h);
    std::string s(which);
    const Crs *c = nullptr;
    if (s == "c2v") c = &m.c2v;
    else if (s == "c2e") c = &m.c2e;
    else if (s == "c2c") c = &m.c2c;
    else if (s == "e2v") c = &m.e2v;
    else if (s == "e2e") c = &m.e2e;
    else if (s == "v2c") c = &m.v2c;
    else if (s == "e2c") c = &m.e2c;
    else if (s == "v2e") c = &m.v2e;
    else throw std::runtime_error("unknown connectivity " + s);
    *off = c->off.data();
    *idx = c->idx.data();
    *rows = c->rows();
  });
}

// FCT setup products: int: maxLevelCell maxLevelEdgeTop nAdvCellsForEdge
// advCellsForEdge highOrderAdvMask; double: edgeSignOnCell meshScalingDel2
// derivTwo advCoef advCoef_3rd
int orc_mesh_get(void *h, const char *name, const void **ptr, int64_t *n, int *is_int) {
  return guard([&] {
    Mesh &m = *static_cast<Mesh *>(h);
    std::string s(name);
    auto ints = [&](std::vector<int> &v) { *ptr = v.data(); *n = int64_t(v.size()); *is_int = 1; };
    auto dbls = [&](std::vector<double> &v) { *ptr = v.data(); *n = int64_t(v.size()); *is_int = 0; };
    if (s == "maxLevelCell") ints(m.maxLevelCell);
    else if (s == "maxLevelEdgeTop") ints(m.maxLevelEdgeTop);
    else if (s == "nAdvCellsForEdge") ints(m.nAdvCellsForEdge);
    else if (s == "advCellsForEdge") ints(m.advCellsForEdge);
    else if (s == "highOrderAdvMask") ints(m.highOrderAdvMask);
    else if (s == "edgeSignOnCell") dbls(m.edgeSignOnCell);
    else if (s == "meshScalingDel2") dbls(m.meshScalingDel2);
    else if (s == "derivTwo") dbls(m.derivTwo);
    else if (s == "advCoef") dbls(m.advCoef);
    else if (s == "advCoef_3rd") dbls(m.advCoef_3rd);
    else throw std::runtime_error("unknown setup field " + s);
  });
}

// ---- individual tasks (AoS reference layouts) ----------------------------
void orc_sw_compute_solve_diagnostics(void *h, int nL, const double *hh, const double *u,
                                      double *h_edge, double *h_vertex, double *circ,
                                      double *vort, double *ke, double *pv_edge,
                                      double *pv_vertex) {
  sw_compute_solve_diagnostics(*static_cast<Mesh *>(h), nL, hh, u, h_edge, h_vertex, circ,
                               vort, ke, pv_edge, pv_vertex);
}
void orc_sw_compute_tend(void *h, int nL, double g, const double *h_edge, const double *hh,
                         const double *h_s, const double *u, const double *ke,
                         const double *pv_edge, double *tend_h, double *tend_u) {
  sw_compute_tend(*static_cast<Mesh *>(h), nL, g, h_edge, hh, h_s, u, ke, pv_edge, tend_h,
                  tend_u);
}
void orc_sw_compute_scalar_tend(void *h, int nL, int nT, const double *h_edge,
                                const double *u, const double *tr, double *tend) {
  sw_compute_scalar_tend(*static_cast<Mesh *>(h), nL, nT, h_edge, u, tr, tend);
}
void orc_sw_del2_viscosity(void *h, int nL, double visc2, const double *u, const double *vort,
                           double *div, double *tend_u) {
  sw_del2_viscosity(*static_cast<Mesh *>(h), nL, visc2, u, vort, div, tend_u);
}
void orc_ocn_calc_normal_thickness_flux(void *h, int nL, const double *lte, const double *nv,
                                        double *ntf) {
  ocn_calc_normal_thickness_flux(*static_cast<Mesh *>(h), nL, lte, nv, ntf);
}
void orc_ocn_calc_provisional_layer_thicknesses(void *h, int nL, const double *lt,
                                                const double *ntf, double *hProv,
                                                double *hProvInv, double *hNewInv, double dt) {
  ocn_calc_provisional_layer_thicknesses(*static_cast<Mesh *>(h), nL, lt, ntf, hProv, hProvInv,
                                         hNewInv, dt);
}
void orc_ocn_low_order(void *h, int nL, int nT, const double *ntf, const double *tr,
                       double *lo) {
  ocn_low_order(*static_cast<Mesh *>(h), nL, nT, ntf, tr, lo);
}
void orc_ocn_high_order(void *h, int nL, int nT, double coef3rd, const double *ntf,
                        const double *tr, double *ho) {
  ocn_high_order(*static_cast<Mesh *>(h), nL, nT, coef3rd, ntf, tr, ho);
}
void orc_ocn_correct_and_rescale(void *h, int nL, int nT, const double *hProvInv,
                                 const double *lt, const double *ntf, const double *lo,
                                 const double *tr, double *ho, double *flxIn, double *flxOut,
                                 double *tend, double *trMax, double *trMin, double dt) {
  ocn_correct_and_rescale(*static_cast<Mesh *>(h), nL, nT, hProvInv, lt, ntf, lo, tr, ho, flxIn,
                          flxOut, tend, trMax, trMin, dt);
}
void orc_ocn_accumulate_fct(void *h, int nL, int nT, const double *hProvInv, const double *lt,
                            const double *ho, const double *tr, double *trProv, double *tend,
                            double dt) {
  ocn_accumulate_fct(*static_cast<Mesh *>(h), nL, nT, hProvInv, lt, ho, tr, trProv, tend, dt);
}
void orc_ocn_hmix_del2(void *h, int nL, int nT, const double *lte, const double *tr,
                       double *tend, double coef) {
  ocn_hmix_del2(*static_cast<Mesh *>(h), nL, nT, lte, tr, tend, coef);
}

// ---- shallow-water driver --------------------------------------------------
void *orc_sw_new(void *mesh, const void *params) {
  try {
    Mesh *m = static_cast<Mesh *>(mesh);
    Params p = to_params(static_cast<const OrcParams *>(params));
    if (!m->fct_ready) setup_fct(*m, p.nL, p.a);
    return new SwRun(m, p);
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return nullptr;
  }
}
void orc_sw_free(void *s) { delete static_cast<SwRun *>(s); }

int orc_sw_set(void *s, const char *name, const double *src, int64_t n) {
  return guard([&] {
    auto *v = sw_field(static_cast<SwRun *>(s), name);
    if (!v || int64_t(v->size()) != n) throw std::runtime_error(std::string("bad field ") + name);
    std::memcpy(v->data(), src, sizeof(double) * n);
  });
}
int orc_sw_get(void *s, const char *name, double *dst, int64_t n) {
  return guard([&] {
    auto *v = sw_field(static_cast<SwRun *>(s), name);
    if (!v || int64_t(v->size()) != n) throw std::runtime_error(std::string("bad field ") + name);
    std::memcpy(dst, v->data(), sizeof(double) * n);
  });
}
// apps/sw/run.cc:88-92: initial diagnostics on curr
int orc_sw_init_diagnostics(void *s) {
  return guard([&] { static_cast<SwRun *>(s)->diagnostics(0); });
}
// apps/sw/run.cc:97-106; *seconds = steady_clock time of the step loop only
int orc_sw_run(void *s, int nsteps, double dt, double *seconds) {
  return guard([&] {
    auto *r = static_cast<SwRun *>(s);
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 1; i <= nsteps; i++) {
      r->advance_timestep(dt);
      r->elapsed += dt;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  });
}
double orc_sw_compute_error(void *s) {
  auto *r = static_cast<SwRun *>(s);
  return sw_compute_error(*r->m, r->p.nL, r->p.a, r->h[0].data(), r->elapsed);
}

// ---- tracer-only ocean driver ----------------------------------------------
void *orc_ocn_new(void *mesh, const void *params, int mono, int hmix, double hmix_coef) {
  try {
    Mesh *m = static_cast<Mesh *>(mesh);
    Params p = to_params(static_cast<const OrcParams *>(params));
    if (!m->fct_ready) setup_fct(*m, p.nL, p.a);
    auto *r = new OcnRun(m, p);
    r->mono = mono;
    r->hmix = hmix;
    r->hmix_coef = hmix_coef;
    return r;
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return nullptr;
  }
}
void orc_ocn_free(void *s) { delete static_cast<OcnRun *>(s); }
// layerThickness (C*nL), normalVelocity (E*nL), tracers (C*nL*nT, curr)
int orc_ocn_init(void *s, const double *lt, const double *nv, const double *tr) {
  return guard([&] {
    auto *r = static_cast<OcnRun *>(s);
    std::memcpy(r->layerThickness[1].data(), lt, sizeof(double) * r->layerThickness[1].size());
    std::memcpy(r->normalVelocity.data(), nv, sizeof(double) * r->normalVelocity.size());
    std::memcpy(r->tr[1].data(), tr, sizeof(double) * r->tr[1].size());
    r->init_layerThicknessEdge();
  });
}
int orc_ocn_run(void *s, int nsteps, double dt, double *seconds) {
  return guard([&] {
    auto *r = static_cast<OcnRun *>(s);
    auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < nsteps; i++) r->rk4_timestep(dt);
    auto t1 = std::chrono::steady_clock::now();
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  });
}
int orc_ocn_get_tracers(void *s, double *dst) {
  return guard([&] {
    auto *r = static_cast<OcnRun *>(s);
    std::memcpy(dst, r->tr[1].data(), sizeof(double) * r->tr[1].size());
  });
}

}  // extern "C"
