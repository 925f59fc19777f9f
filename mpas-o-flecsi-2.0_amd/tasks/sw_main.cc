// `sw`: the reference's shallow-water executable (apps/sw/sw.cc,
// initialize.cc, run.cc, finalize.cc) on a generated mesh, through the C++
// task layer and the MI355X library.
//
//   sw [--level L] [--unrotated] [-t 1|5|balanced] [-s nsteps] [-d dt]
//      [--tracers N] [--fused] [--num_output N] [--output PATH]
//
// Defaults follow apps/sw/state.hh:51-64 (test case 1, 10 steps, dt 40 s).
// Case 1 ends with the reference's L2 gate (apps/sw/finalize.cc:15-29):
// exit status 1 when the error exceeds 0.0015.  One JSON line reports the
// error and an FNV-1a hash of the final h bytes.  With --num_output N (the
// reference's option, mpasoflecsi/io/state.hh:13-17) the run writes its
// output file (default output.ncdf) as apps/sw/run.cc:73-113 does.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../host/initial.h"
#include "sw_run.hh"
#ifdef MPAS_WITH_IO
#include "mpas_io.h"
#endif

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

#ifdef MPAS_WITH_IO
// The output file of apps/sw/run.cc:73-113: mesh_output_init, then the
// fields of apps/sw/state.hh:66-70 per output time.
class output_file {
 public:
  output_file(const std::string &path, int ntimes, const mg_mesh *g, int64_t nTracers) {
    auto D = [&](const char *n) { return mg_dbl_field(g, n, nullptr); };
    mpas_output_mesh om{};
    om.nCells = mg_dim(g, "nCells");
    om.nEdges = mg_dim(g, "nEdges");
    om.nVertices = nV_ = mg_dim(g, "nVertices");
    om.vertexDegree = 3;
    om.nVertLevels = 1;
    om.nTracers = nTracers;
    om.areaCell = D("areaCell");
    om.xCell = D("xCell");
    om.yCell = D("yCell");
    om.zCell = D("zCell");
    om.latCell = D("latCell");
    om.lonCell = D("lonCell");
    om.latVertex = D("latVertex");
    om.lonVertex = D("lonVertex");
    om.meshDensity = D("meshDensity");
    om.dvEdge = D("dvEdge");
    om.dcEdge = D("dcEdge");
    om.cellsOnVertex = mg_int_field(g, "cellsOnVertex", nullptr);
    if (int rc = mpas_output_create(path.c_str(), ntimes, &om, &o_))
      throw error(rc, std::string("mpas_output_create: ") + mpas_io_last_error());
    nT_ = nTracers;
  }
  ~output_file() {
    if (o_) mpas_output_close(o_);
  }
  // The five fields of one output time; vorticity and pv_vertex are zero at
  // time 0 (init_extra_fields, apps/sw/tasks/testcase.cc:16-35).
  void write(int t, const std::vector<double> &h, const std::vector<double> &h_s,
             const std::vector<double> &vort, const std::vector<double> &pv,
             const std::vector<double> &tr) {
    std::vector<double> zero(size_t(nV_), 0.0);
    const double *v = t == 0 ? zero.data() : vort.data(), *p = t == 0 ? zero.data() : pv.data();
    const mpas_output_field fs[] = {{"h", h.data(), MPAS_OUT_CELLS, 1, 0},
                                    {"h_s", h_s.data(), MPAS_OUT_CELLS, 0, 0},
                                    {"vorticity", v, MPAS_OUT_VERTICES, 1, 0},
                                    {"potential_vorticity", p, MPAS_OUT_VERTICES, 1, 0},
                                    {"tracers", tr.data(), MPAS_OUT_CELLS, 1, int32_t(nT_)}};
    for (const auto &x : fs) {
      if (std::string(x.name) == "tracers" && !nT_) continue;
      if (int rc = mpas_output_write(o_, t, &x))
        throw error(rc, std::string("mpas_output_write: ") + mpas_io_last_error());
    }
  }

 private:
  mpas_output *o_ = nullptr;
  int64_t nV_ = 0, nT_ = 0;
};
#endif

}  // namespace

int main(int argc, char **argv) {
  int level = 5, nsteps = 10, ntr = 0, rotate = 1;
  double dt = 40;
  std::string tc = "1";
  bool fused = false;
  int num_output = 0;
  std::string out_path = "output.ncdf";
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
    else if (a == "--num_output") num_output = std::atoi(next());
    else if (a == "--output") out_path = next();
    else {
      std::fprintf(stderr, "usage: sw [--level L] [--unrotated] [-t 1|5|balanced] [-s n] "
                           "[-d dt] [--tracers N] [--fused] [--num_output N] [--output PATH]\n");
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
#ifndef MPAS_WITH_IO
  if (num_output > 0) {
    std::fprintf(stderr, "--num_output needs the HDF5 output library (lib/libmpas_io.so)\n");
    mg_free(g);
    return 2;
  }
#endif
  try {
    mesh m(d, p);
#ifdef MPAS_WITH_IO
    std::unique_ptr<output_file> out;
    if (num_output > 0) out = std::make_unique<output_file>(out_path, num_output, g, ntr);
#endif
    std::vector<double> h_final;
    if (fused) {
      check(mpas_sw_state_upload(m.ctx(), h.data(), u.data(), ntr ? tr.data() : nullptr,
                                 h_s.data()),
            "mpas_sw_state_upload");
#ifdef MPAS_WITH_IO
      // apps/sw/run.cc:77-106 with the steps between output times fused.
      auto field_of = [&](const char *name, int64_t n) {
        std::vector<double> v(static_cast<size_t>(n));
        if (n) check(mpas_sw_field_download(m.ctx(), name, v.data(), n), "mpas_sw_field_download");
        return v;
      };
      auto write = [&](int t) {
        out->write(t, field_of("h", nC), field_of("h_s", nC), field_of("vorticity", nV),
                   field_of("pv_vertex", nV), field_of("tracers", nC * ntr));
      };
      int done = 0, t = 0;
      if (out) {
        write(t++);
        for (int i : sw::output_steps(nsteps, num_output)) {
          check(mpas_rk4_step(m.ctx(), dt, i - done), "mpas_rk4_step");
          done = i;
          write(t++);
        }
      }
      if (nsteps > done) check(mpas_rk4_step(m.ctx(), dt, nsteps - done), "mpas_rk4_step");
#else
      check(mpas_rk4_step(m.ctx(), dt, nsteps), "mpas_rk4_step");
#endif
      h_final.resize(nC);
      check(mpas_sw_state_download(m.ctx(), h_final.data(), nullptr, nullptr),
            "mpas_sw_state_download");
    } else {
      sw::fields f(m);
      f.h[sw::curr].upload(h);
      f.u[sw::curr].upload(u);
      if (ntr) f.tracers[sw::curr].upload(tr);
      f.bottomDepth.upload(h_s);
#ifdef MPAS_WITH_IO
      if (out) {
        sw::run(m, f, dt, nsteps, num_output, [&](int t) {
          out->write(t, f.h[sw::curr].download(), f.bottomDepth.download(),
                     f.vorticity.download(), f.pv_vertex.download(),
                     ntr ? f.tracers[sw::curr].download() : std::vector<double>());
        });
      } else {
        sw::run(m, f, dt, nsteps);
      }
#else
      sw::run(m, f, dt, nsteps);
#endif
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
