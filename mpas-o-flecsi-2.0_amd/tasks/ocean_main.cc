// `ocean`: the reference's tracer-only ocean executable (apps/ocean:
// initialize.cc:59-104, run.cc:10-76, finalize.cc:14-31) on a generated
// mesh, through the fused ocean step of the MI355X library.
//
//   ocean [--level L] [-t 1|2|3] [-s nsteps] [-d dt]
//
// Defaults follow apps/ocean/state.hh:51-64 (test case 1, 1 step, dt 96 s).
// The test cases keep the reference's set-up (initialize.cc:59-84):
//   1  Gaussian hill in solid-body rotation, monotone FCT advection;
//   2  falls through into case 3 (no break, initialize.cc:71-76): the
//      case-3 state with FCT advection and del2 mixing (1e7 m^2/s), scored
//      against the slotted cylinder by finalize, as the reference does;
//   3  spherical harmonic at rest, del2 mixing (1e7 m^2/s) only.
// The finalize report is calc_error_case_<t> (testcase.cc:175-294). One
// JSON line gives it with an FNV-1a hash of the final tracer bytes.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../host/initial.h"
#include "../host/meshgen.h"
#include "mpas_hip.h"

namespace {

constexpr double kA = 6371000.0;  // mpas_constants::a

uint64_t fnv1a(const std::vector<double> &v) {
  uint64_t h = 1469598103934665603ull;
  const unsigned char *p = reinterpret_cast<const unsigned char *>(v.data());
  for (size_t i = 0; i < v.size() * sizeof(double); i++) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

int fail(const char *what) {
  std::fprintf(stderr, "error: %s: %s\n", what, mpas_last_error());
  return 3;
}

}  // namespace

int main(int argc, char **argv) {
  int level = 5, nsteps = 1, tc = 1;
  double dt = 96.0;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    auto next = [&]() -> const char * {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "missing value for %s\n", a.c_str());
        std::exit(2);
      }
      return argv[++i];
    };
    if (a == "--level") level = std::atoi(next());
    else if (a == "-t" || a == "--testcase") tc = std::atoi(next());
    else if (a == "-s" || a == "--nsteps") nsteps = std::atoi(next());
    else if (a == "-d" || a == "--dt") dt = std::atof(next());
    else {
      std::fprintf(stderr, "usage: ocean [--level L] [-t 1|2|3] [-s n] [-d dt]\n");
      return 2;
    }
  }
  if (tc < 1 || tc > 3 || nsteps < 0) {
    std::fprintf(stderr, "test case must be 1, 2 or 3 and nsteps >= 0\n");
    return 2;
  }
  mg_mesh *g = mg_generate_oriented(level, kA, 2, 1, 7.29212e-5, 1);
  if (!g) {
    std::fprintf(stderr, "mesh generation failed: %s\n", mg_last_error());
    return 2;
  }
  const int64_t nC = mg_dim(g, "nCells"), nE = mg_dim(g, "nEdges"), nV = mg_dim(g, "nVertices");
  auto I = [&](const char *n) { return mg_int_field(g, n, nullptr); };
  auto D = [&](const char *n) { return mg_dbl_field(g, n, nullptr); };
  mpas_mesh_desc d{};
  d.nCells = int32_t(nC);
  d.nEdges = int32_t(nE);
  d.nVertices = int32_t(nV);
  d.maxEdges = 8;
  d.maxEdges2 = 16;
  d.vertexDegree = 3;
  d.verticesOnCell = I("verticesOnCell");
  d.edgesOnCell = I("edgesOnCell");
  d.cellsOnCell = I("cellsOnCell");
  d.verticesOnEdge = I("verticesOnEdge");
  d.edgesOnEdge = I("edgesOnEdge");
  d.areaCell = D("areaCell");
  d.dvEdge = D("dvEdge");
  d.dcEdge = D("dcEdge");
  d.areaTriangle = D("areaTriangle");
  d.kiteAreasOnVertex = D("kiteAreasOnVertex");
  d.weightsOnEdge = D("weightsOnEdge");
  d.fVertex = D("fVertex");
  d.xCell = D("xCell");
  d.yCell = D("yCell");
  d.zCell = D("zCell");
  d.xVertex = D("xVertex");
  d.yVertex = D("yVertex");
  d.zVertex = D("zVertex");
  d.meshDensity = D("meshDensity");

  // initialize.cc:59-84: case 2 runs the case-3 set-up with advection on
  const double hmix = tc == 1 ? 0.0 : 1e7;
  const int setup = tc == 1 ? 1 : 3;
  std::vector<double> nv(static_cast<size_t>(nE)), lt(static_cast<size_t>(nC)),
      tr(static_cast<size_t>(nC));
  mg_init_ocean_case(g, setup, kA, hmix, 0.0, nv.data(), lt.data(), tr.data());

  // Case 3 has advection off: it runs the same FCT kernels at zero velocity,
  // where every flux is +-0 and the tendency is the mixing term alone.
  mpas_params p{};
  p.nLevels = 1;
  p.nTracers = 1;
  p.tracer_scheme = MPAS_TRACER_FCT;
  p.gravity = 9.80616;
  p.sphere_radius = kA;
  p.coef3rd = 0.25;
  p.kappa2 = hmix;
  mpas_ctx *ctx = nullptr;
  if (mpas_ctx_create(&d, &p, &ctx)) {
    mg_free(g);
    return fail("mpas_ctx_create");
  }
  std::vector<double> h_s(static_cast<size_t>(nC), 0.0);
  int rc = 0;
  if (mpas_sw_state_upload(ctx, lt.data(), nv.data(), tr.data(), h_s.data()))
    rc = fail("mpas_sw_state_upload");
  else if (nsteps > 0 && mpas_ocean_rk4_step(ctx, dt, nsteps))
    rc = fail("mpas_ocean_rk4_step");
  else if (mpas_sw_state_download(ctx, nullptr, nullptr, tr.data()))
    rc = fail("mpas_sw_state_download");
  if (!rc) {
    // finalize.cc:14-31 with the elapsed time summed as run.cc:65-72 sums it
    double elapsed = 0.0;
    for (int i = 0; i < nsteps; i++) elapsed += dt;
    double l2 = 0, mx = 0;
    mg_ocean_error(g, tc, kA, hmix, elapsed, tr.data(), 1, &l2, &mx);
    std::printf("{\"testcase\": %d, \"level\": %d, \"cells\": %lld, \"steps\": %d, \"dt\": %g, "
                "\"l2_error\": %.17g, \"max_error\": %.17g, \"tracer_fnv1a\": \"%016llx\"}\n",
                tc, level, (long long)nC, nsteps, dt, l2, mx, (unsigned long long)fnv1a(tr));
  }
  mpas_ctx_destroy(ctx);
  mg_free(g);
  return rc;
}
