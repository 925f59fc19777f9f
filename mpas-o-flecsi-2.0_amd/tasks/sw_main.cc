// `sw`: the reference's shallow-water executable (apps/sw/sw.cc,
// initialize.cc, run.cc, finalize.cc) on a generated mesh, through the C++
// task layer and the MI355X library.
//
//   sw [--level L] [--unrotated] [-t 1|5|balanced] [-s nsteps] [-d dt]
//      [--tracers N] [--fused]
//
// Defaults follow apps/sw/state.hh:51-64 (test case 1, 10 steps, dt 40 s).
// Case 1 ends with the reference's L2 gate (apps/sw/finalize.cc:15-29):
// exit status 1 when the error exceeds 0.0015.  One JSON line reports the
// error and an FNV-1a hash of the final h bytes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../host/initial.h"
#include "sw_run.hh"

using namespace mpas::hip;

namespace {

constexpr double kA = 6371000.0;      // mpas_constants::a
constexpr double kOmega = 7.29212e-5;  // mpas_constants::omega
constexpr double kG = 9.80616;         // mpas_constants::gravity

std::vector<double> dbl(const mg_mesh *m, const char *n) {
  int64_t k = 0;
  const double *p = mg_dbl_field(m, n, &k);
  return std::vector<double>(p, p + k);
}

uint64_t fnv1a(const std::vector<double> &v) {
  uint64_t h = 1469598103934665603ull;
  const unsigned char *p = reinterpret_cast<const unsigned char *>(v.data());
  for (size_t i = 0; i < v.size() * sizeof(double); i++) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

}  // namespace

int main(int argc, char **argv) {
  int level = 5, nsteps = 10, ntr = 0, rotate = 1;
  double dt = 40;
  std::string tc = "1";
  bool fused = false;
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
    else if (a == "--unrotated") rotate = 0;
    else if (a == "-t" || a == "--testcase") tc = next();
    else if (a == "-s" || a == "--nsteps") nsteps = std::atoi(next());
    else if (a == "-d" || a == "--dt") dt = std::atof(next());
    else if (a == "--tracers") ntr = std::atoi(next());
    else if (a == "--fused") fused = true;
    else {
      std::fprintf(stderr, "usage: sw [--level L] [--unrotated] [-t 1|5|balanced] [-s n] "
                           "[-d dt] [--tracers N] [--fused]\n");
      return 2;
    }
  }
  const double radius = tc == "balanced" ? 6371220.0 : kA;
  mg_mesh *g = mg_generate_oriented(level, radius, 2, 1, tc == "balanced" ? 7.292e-5 : kOmega,
                                    rotate);
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
  d.xCell = D("xCell");
  d.yCell = D("yCell");
  d.zCell = D("zCell");
  d.xVertex = D("xVertex");
  d.yVertex = D("yVertex");
  d.zVertex = D("zVertex");
  d.meshDensity = D("meshDensity");

  std::vector<double> u(nE), h(nC), h_s(nC, 4500.0), fV = dbl(g, "fVertex");
  std::vector<double> tr(size_t(nC) * (ntr > 0 ? ntr : 0), 0.0);
  if (tc == "1") {
    mg_init_sw_case1(g, kA, u.data(), h.data());
  } else if (tc == "5") {
    std::vector<double> t5(size_t(nC) * std::max(ntr, 1));
    mg_init_sw_case5(g, kA, kOmega, kG, u.data(), h.data(), h_s.data(), fV.data(), t5.data(),
                     std::max(ntr, 1));
    if (ntr) tr = t5;
  } else if (tc == "balanced") {
    mg_init_sw_balanced(g, radius, 7.292e-5, kG, 20.0, 5000.0, 10.0, 500e3, 1, u.data(),
                        h.data(), nullptr);
    for (int t = 0; t < ntr; t++) {
      std::vector<double> one(nC);
      mg_init_tracer_blobs(g, radius, 8, 300e3, 1500e3, 2 + t, one.data());
      for (int64_t c = 0; c < nC; c++) tr[c * ntr + t] = one[c];
    }
  } else {
    std::fprintf(stderr, "unknown test case %s\n", tc.c_str());
    return 2;
  }
  d.fVertex = fV.data();

  mpas_params p{};
  p.nLevels = 1;
  p.nTracers = ntr;
  p.tracer_scheme = ntr ? MPAS_TRACER_CENTRED : MPAS_TRACER_NONE;
  p.freeze_u = tc == "1";
  p.gravity = kG;
  p.sphere_radius = radius;
  p.coef3rd = 0.25;
  try {
    mesh m(d, p);
    std::vector<double> h_final;
    if (fused) {
      check(mpas_sw_state_upload(m.ctx(), h.data(), u.data(), ntr ? tr.data() : nullptr,
                                 h_s.data()),
            "mpas_sw_state_upload");
      check(mpas_rk4_step(m.ctx(), dt, nsteps), "mpas_rk4_step");
      h_final.resize(nC);
      check(mpas_sw_state_download(m.ctx(), h_final.data(), nullptr, nullptr),
            "mpas_sw_state_download");
    } else {
      sw::fields f(m);
      f.h[sw::curr].upload(h);
      f.u[sw::curr].upload(u);
      if (ntr) f.tracers[sw::curr].upload(tr);
      f.bottomDepth.upload(h_s);
      sw::run(m, f, dt, nsteps);
      m.sync();
      h_final = f.h[sw::curr].download();
    }
    double err = -1;
    if (tc == "1") {
      field hf(m, entity::cells, 1), lat(m, entity::cells, 1), lon(m, entity::cells, 1);
      hf.upload(h_final);
      lat.upload(dbl(g, "latCell"));
      lon.upload(dbl(g, "lonCell"));
      err = sw::task::compute_error(m, hf, lat, lon, dt * nsteps);
    }
    std::printf("{\"testcase\": \"%s\", \"level\": %d, \"cells\": %lld, \"steps\": %d, "
                "\"dt\": %g, \"driver\": \"%s\", \"l2_error\": %.17g, \"h_fnv1a\": "
                "\"%016llx\"}\n",
                tc.c_str(), level, (long long)nC, nsteps, dt, fused ? "fused" : "tasks", err,
                (unsigned long long)fnv1a(h_final));
    mg_free(g);
    if (tc == "1" && err > 0.0015) {
      std::fprintf(stderr, "case 1 L2 error %.6g exceeds 0.0015\n", err);
      return 1;
    }
  } catch (const error &e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    mg_free(g);
    return 3;
  }
  return 0;
}
