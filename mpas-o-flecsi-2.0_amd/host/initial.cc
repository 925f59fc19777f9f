// Initial states for the shallow-water / tracer runs.
//
// Reference test cases (restated so that the reference's own checks can run
// on generated meshes):
//   sw case 1  apps/sw/tasks/testcase.cc:37-77   (cosine bell, frozen wind)
//   sw case 5  apps/sw/tasks/testcase.cc:80-168  (zonal flow over a mountain)
//   sw case 6  apps/sw/tasks/testcase.cc:171-227 (Rossby-Haurwitz wave)
//   ocean 1-3  apps/ocean/tasks/testcase.cc:16-173
// Seeded states of the benchmark workloads (no reference counterpart; the
// construction follows setup_case_5 without the mountain):
//   balanced solid-body zonal flow + a Gaussian thickness bump,
//   tracers as normalised sums of Gaussian blobs.
// Random numbers come from SplitMix64 (Steele, Lea, Flood 2014), so every
// workload is reproducible from its seed on any machine.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "initial.h"

namespace {

const double kPi = 3.141592653589793;

struct SplitMix64 {
  uint64_t x;
  uint64_t next() {
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * 0x1.0p-53; }
  double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
};

double great_circle(double lat1, double lon1, double lat2, double lon2) {
  double s1 = std::sin(0.5 * (lat2 - lat1)), s2 = std::sin(0.5 * (lon2 - lon1));
  double arg = std::sqrt(s1 * s1 + std::cos(lat1) * std::cos(lat2) * s2 * s2);
  return 2.0 * std::asin(std::min(1.0, arg));
}

const double *F(const mg_mesh *m, const char *n) { return mg_dbl_field(m, n, nullptr); }
const int32_t *I(const mg_mesh *m, const char *n) { return mg_int_field(m, n, nullptr); }

// u = -(psi(v2) - psi(v1)) / dvEdge with psi(v) = -a u0 sin(latVertex)
// (apps/sw/tasks/testcase.cc:104-117, alpha = 0)
void zonal_wind(const mg_mesh *m, double a, double u0, double *u) {
  const int64_t nE = mg_dim(m, "nEdges");
  const double *latV = F(m, "latVertex"), *dv = F(m, "dvEdge");
  const int32_t *voe = I(m, "verticesOnEdge");
  for (int64_t e = 0; e < nE; e++) {
    int v1 = voe[2 * e] - 1, v2 = voe[2 * e + 1] - 1;
    double p1 = -a * u0 * std::sin(latV[v1]);
    double p2 = -a * u0 * std::sin(latV[v2]);
    u[e] = -1.0 * (p2 - p1) / dv[e];
  }
}

}  // namespace

extern "C" {

// Balanced zonal flow u = u0 cos(lat) with the geostrophic thickness
//   h = (g h0 - (a omega u0 + u0^2/2) sin^2(lat)) / g,
// h0 chosen so the area mean of h is h_mean, plus a Gaussian bump of
// amplitude bump_amp and e-folding radius bump_radius (metres) centred at a
// point drawn uniformly on the sphere from `seed`.  u: [nEdges], h: [nCells].
// Returns the bump centre in centre_latlon[2] when non-null.
int mg_init_sw_balanced(const mg_mesh *m, double a, double omega, double gravity,
                        double u0, double h_mean, double bump_amp,
                        double bump_radius, uint64_t seed, double *u, double *h,
                        double *centre_latlon) {
  if (!m || !u || !h) return -1;
  const int64_t nC = mg_dim(m, "nCells");
  const double *latC = F(m, "latCell"), *lonC = F(m, "lonCell");
  zonal_wind(m, a, u0, u);
  SplitMix64 rng{seed};
  double lat_c = std::asin(rng.uniform(-1.0, 1.0));
  double lon_c = rng.uniform(0.0, 2.0 * kPi);
  if (centre_latlon) {
    centre_latlon[0] = lat_c;
    centre_latlon[1] = lon_c;
  }
  const double K = a * omega * u0 + 0.5 * u0 * u0;
  const double h0 = h_mean + K / (3.0 * gravity);
  for (int64_t c = 0; c < nC; c++) {
    double s = std::sin(latC[c]);
    double hb = (gravity * h0 - K * s * s) / gravity;
    double r = a * great_circle(lat_c, lon_c, latC[c], lonC[c]);
    h[c] = hb + bump_amp * std::exp(-(r / bump_radius) * (r / bump_radius));
  }
  return 0;
}

// Tracer field in [0,1]: sum of `nblobs` Gaussian blobs (centre uniform on
// the sphere, amplitude U[0.5,1], radius U[rmin,rmax] metres) divided by its
// maximum over the cells.  tr: [nCells].
int mg_init_tracer_blobs(const mg_mesh *m, double a, int nblobs, double rmin,
                         double rmax, uint64_t seed, double *tr) {
  if (!m || !tr || nblobs < 1) return -1;
  const int64_t nC = mg_dim(m, "nCells");
  const double *latC = F(m, "latCell"), *lonC = F(m, "lonCell");
  SplitMix64 rng{seed};
  std::vector<double> bl(nblobs), bn(nblobs), ba(nblobs), br(nblobs);
  for (int k = 0; k < nblobs; k++) {
    bl[k] = std::asin(rng.uniform(-1.0, 1.0));
    bn[k] = rng.uniform(0.0, 2.0 * kPi);
    ba[k] = rng.uniform(0.5, 1.0);
    br[k] = rng.uniform(rmin, rmax);
  }
  double mx = 0.0;
  for (int64_t c = 0; c < nC; c++) {
    double s = 0.0;
    for (int k = 0; k < nblobs; k++) {
      double r = a * great_circle(bl[k], bn[k], latC[c], lonC[c]) / br[k];
      s += ba[k] * std::exp(-r * r);
    }
    tr[c] = s;
    mx = std::max(mx, s);
  }
  for (int64_t c = 0; c < nC; c++) tr[c] /= mx;
  return 0;
}

// apps/sw/tasks/testcase.cc:37-77 (case 1: cosine bell, alpha = 0)
int mg_init_sw_case1(const mg_mesh *m, double a, double *u, double *h) {
  if (!m || !u || !h) return -1;
  const int64_t nC = mg_dim(m, "nCells");
  const double *latC = F(m, "latCell"), *lonC = F(m, "lonCell");
  const double u0 = 2.0 * kPi * a / (12.0 * 86400.0);
  const double h0 = 1000.0, theta_c = 0.0, lambda_c = 3.0 * kPi / 2.0;
  zonal_wind(m, a, u0, u);
  for (int64_t c = 0; c < nC; c++) {
    double s1 = std::sin(0.5 * (latC[c] - theta_c)), s2 = std::sin(0.5 * (lonC[c] - lambda_c));
    double r = 2 * a * std::asin(std::sqrt(s1 * s1 + std::cos(theta_c) * std::cos(latC[c]) * (s2 * s2)));
    h[c] = (r < a / 3.0) ? (h0 / 2.0) * (1.0 + std::cos(kPi * r * 3.0 / a)) : h0 / 2.0;
  }
  return 0;
}

// apps/sw/tasks/testcase.cc:80-168 (case 5: h_s mountain, tracers 0 and 1)
// lonCell is shifted into [0, 2 pi) there (line 134-135); generated meshes
// already use that range.  tr: [nCells][ntr] with ntr >= 1.
int mg_init_sw_case5(const mg_mesh *m, double a, double omega, double gravity,
                     double *u, double *h, double *h_s, double *fVertex,
                     double *tr, int ntr) {
  if (!m || !u || !h || !h_s) return -1;
  const int64_t nC = mg_dim(m, "nCells"), nV = mg_dim(m, "nVertices");
  const double *latC = F(m, "latCell"), *lonC = F(m, "lonCell"), *latV = F(m, "latVertex");
  const double u0 = 20, gh0 = 5960.0 * gravity, hs0 = 2000;
  const double theta_c = kPi / 6.0, lambda_c = 3.0 * kPi / 2.0, rr = kPi / 9.0;
  zonal_wind(m, a, u0, u);
  if (fVertex)
    for (int64_t v = 0; v < nV; v++) fVertex[v] = 2.0 * omega * std::sin(latV[v]);
  for (int64_t c = 0; c < nC; c++) {
    double dl = lonC[c] - lambda_c, dt = latC[c] - theta_c;
    double r = std::sqrt(std::min(rr * rr, dl * dl + dt * dt));
    h_s[c] = hs0 * (1.0 - r / rr);
    if (tr && ntr > 0) tr[c * ntr + 0] = 1.0 - r / rr;
    if (tr && ntr > 1) {
      double dt2 = latC[c] - theta_c - kPi / 6.0;
      double r2 = std::sqrt(std::min(rr * rr, dl * dl + dt2 * dt2));
      tr[c * ntr + 1] = 1.0 - r2 / rr;
    }
    double s = std::sin(latC[c]);
    h[c] = (gh0 - (a * omega * u0 + 0.5 * u0 * u0) * s * s) / gravity - h_s[c];
  }
  return 0;
}

// apps/sw/tasks/testcase.cc:171-227 (case 6: Rossby-Haurwitz wave, R = 4)
int mg_init_sw_case6(const mg_mesh *m, double a, double omega, double gravity,
                     double *u, double *h) {
  if (!m || !u || !h) return -1;
  const int64_t nC = mg_dim(m, "nCells"), nE = mg_dim(m, "nEdges");
  const double *latC = F(m, "latCell"), *lonC = F(m, "lonCell");
  const double *latV = F(m, "latVertex"), *lonV = F(m, "lonVertex"), *dv = F(m, "dvEdge");
  const int32_t *voe = I(m, "verticesOnEdge");
  const double h0 = 8000.0, w = 7.848e-6, K = 7.848e-6, R = 4.0;
  auto psi = [&](int v) {
    return -a * a * w * std::sin(latV[v]) +
           a * a * K * std::pow(std::cos(latV[v]), R) * std::sin(latV[v]) * std::cos(R * lonV[v]);
  };
  for (int64_t e = 0; e < nE; e++) {
    int v1 = voe[2 * e] - 1, v2 = voe[2 * e + 1] - 1;
    u[e] = -1.0 * (psi(v2) - psi(v1)) / dv[e];
  }
  for (int64_t c = 0; c < nC; c++) {
    double th = latC[c], ct = std::cos(th);
    double aa = 0.5 * w * (2.0 * omega + w) * ct * ct +
                0.25 * K * K * std::pow(ct, 2 * R) *
                    ((R + 1.0) * ct * ct + 2.0 * R * R - R - 2.0 - 2.0 * R * R / (ct * ct));
    double bb = (2.0 * (omega + w) * K / ((R + 1.0) * (R + 2.0))) * std::pow(ct, R) *
                ((R * R + 2.0 * R + 2.0) - (R + 1.0) * ct * (R + 1.0) * ct);
    double cc = 0.25 * K * K * std::pow(ct, 2.0 * R) * ((R + 1.0) * ct * ct - R - 2.0);
    h[c] = (gravity * h0 + a * a * aa + a * a * bb * std::cos(R * lonC[c]) +
            a * a * cc * std::cos(2.0 * R * lonC[c])) /
           gravity;
  }
  return 0;
}

// apps/ocean/tasks/testcase.cc:16-173.  which = 1 Gaussian hill,
// 2 slotted cylinder, 3 spherical harmonic (coef = hmix coefficient).
// normalVelocity: [nEdges], layerThickness: [nCells] (= 1), tr: [nCells].
int mg_init_ocean_case(const mg_mesh *m, int which, double a, double coef,
                       double elapsed, double *normalVelocity,
                       double *layerThickness, double *tr) {
  if (!m) return -1;
  const int64_t nC = mg_dim(m, "nCells"), nE = mg_dim(m, "nEdges");
  const double *latE = F(m, "latEdge"), *angE = F(m, "angleEdge");
  const double *latC = F(m, "latCell"), *lonC = F(m, "lonCell");
  const double *xC = F(m, "xCell"), *yC = F(m, "yCell"), *zC = F(m, "zCell");
  const double u0 = 2.0 * kPi * a / 86400.0;
  for (int64_t e = 0; e < nE; e++)
    normalVelocity[e] = (which == 3) ? 0.0 : u0 * std::cos(latE[e]) * std::cos(angE[e]);
  for (int64_t c = 0; c < nC; c++) layerThickness[c] = 1.;
  for (int64_t c = 0; c < nC; c++) {
    if (which == 1) {
      const double maxG = 21983.7698959881, Knorm = 20.0 / maxG, kappa = 10.0, T0 = 50.0;
      // x' = x cos(-rot) - y sin(-rot), rot = 2 pi (elapsed / 86400)
      // (testcase.cc:188-197)
      double rot = 2.0 * kPi * (elapsed / 86400.);
      double xp = xC[c] * std::cos(-rot) - yC[c] * std::sin(-rot);
      tr[c] = T0 + Knorm * std::exp(kappa * xp / a);
    } else if (which == 2) {
      double rot = 2.0 * kPi * (elapsed / 86400.);  // testcase.cc:236-245
      double x = xC[c] * std::cos(rot) + yC[c] * std::sin(rot);
      double y = -xC[c] * std::sin(rot) + yC[c] * std::cos(rot);
      double lon = lonC[c] - rot;
      double aa = a / 2., hw = a / 12., bot = -a / 3.;
      double cx = a - x, cy = 0. - y, cz = 0. - zC[c];
      double rr = std::sqrt(x * x + y * y + zC[c] * zC[c]);
      double dist = rr * 2. * std::asin(std::sqrt(cx * cx + cy * cy + cz * cz) / (2. * rr));
      double longC = lon;
      if (longC > kPi) longC -= 2. * kPi;
      if (dist > aa) tr[c] = 50.;
      else if ((a * latC[c] > bot) && (a * latC[c] < aa) && (a * std::abs(longC) < hw)) tr[c] = 50.;
      else tr[c] = 100.;
    } else {
      double cl = std::cos(latC[c]);
      tr[c] = 50. + 20. * cl * cl * std::cos(2. * lonC[c]) * std::exp(-6.0 * coef * elapsed / (a * a));
    }
  }
  return 0;
}

int mg_ocean_error(const mg_mesh *m, int which, double a, double coef, double elapsed,
                   const double *tr, int64_t stride, double *l2, double *max_err) {
  if (!m || !tr || !l2 || !max_err || stride < 1 || which < 1 || which > 3) return -1;
  const int64_t nC = mg_dim(m, "nCells"), nE = mg_dim(m, "nEdges");
  std::vector<double> nv(static_cast<size_t>(nE)), lt(static_cast<size_t>(nC)),
      truth(static_cast<size_t>(nC));
  if (mg_init_ocean_case(m, which, a, coef, elapsed, nv.data(), lt.data(), truth.data()))
    return -1;
  double errSqd = 0.0, errMax = 0.0;
  for (int64_t c = 0; c < nC; c++) {
    const double error = (tr[c * stride] - truth[c]) / truth[c];
    if (std::abs(error) > errMax) errMax = error;
    errSqd += error * error;
  }
  *l2 = std::sqrt(errSqd) / double(nC);
  *max_err = std::abs(errMax);
  return 0;
}

}  // extern "C"
