// ORACLE — test infrastructure only (see orc_internal.h).
// One-time FCT / del2 setup restated from apps/ocean/tasks/init_functions.cc
// with the geometry helpers of mpasoflecsi/eqns/metrics.hh and the dense
// linear algebra of mpasoflecsi/eqns/matrix.hh.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <vector>

#include "orc_internal.h"

namespace orc {
namespace {

using Mat = std::vector<std::vector<double>>;

// mpasoflecsi/eqns/metrics.hh:23-73
double sphere_angle(double ax, double ay, double az, double bx, double by,
                    double bz, double cx, double cy, double cz) {
  double a, b, c;
  a = std::acos(std::max(std::min(bx * cx + by * cy + bz * cz, 1.0), -1.0));
  b = std::acos(std::max(std::min(ax * cx + ay * cy + az * cz, 1.0), -1.0));
  c = std::acos(std::max(std::min(ax * bx + ay * by + az * bz, 1.0), -1.0));
  double ABx = bx - ax, ABy = by - ay, ABz = bz - az;
  double ACx = cx - ax, ACy = cy - ay, ACz = cz - az;
  double Dx = (ABy * ACz) - (ABz * ACy);
  double Dy = (ABz * ACx) - (ABx * ACz);
  double Dz = (ABx * ACy) - (ABy * ACx);
  double s = 0.5 * (a + b + c);
  double sin_angle =
      std::sqrt(std::min(1.0, std::max(0.0, (std::sin(s - b) * std::sin(s - c)) /
                                                (std::sin(b) * std::sin(c)))));
  if (Dx * ax + Dy * ay + Dz * az >= 0.0)
    return 2.0 * std::asin(std::max(std::min(sin_angle, 1.0), -1.0));
  return -2.0 * std::asin(std::max(std::min(sin_angle, 1.0), -1.0));
}

// mpasoflecsi/eqns/metrics.hh:75-91
double sphere_arc_length(double ax, double ay, double az, double bx, double by,
                         double bz) {
  double cx = bx - ax, cy = by - ay, cz = bz - az;
  double r = std::sqrt(ax * ax + ay * ay + az * az);
  double c = std::sqrt(cx * cx + cy * cy + cz * cz);
  return r * 2. * std::asin(c / (2. * r));
}

// mpasoflecsi/eqns/metrics.hh:93-111
void arc_bisect(double ax, double ay, double az, double bx, double by, double bz,
                double &cx, double &cy, double &cz) {
  double r = std::sqrt(ax * ax + ay * ay + az * az);
  cx = 0.5 * (ax + bx);
  cy = 0.5 * (ay + by);
  cz = 0.5 * (az + bz);
  if (!(cx != 0. && cy != 0. && cz != 0.))
    throw std::runtime_error("arc_bisect: A and B are diametrically opposite");
  double d = std::sqrt(cx * cx + cy * cy + cz * cz);
  cx = r * cx / d;
  cy = r * cy / d;
  cz = r * cz / d;
}

// mpasoflecsi/eqns/matrix.hh:27-37
Mat transpose(const Mat &in) {
  Mat out(in[0].size(), std::vector<double>(in.size()));
  for (size_t i = 0; i < in.size(); i++)
    for (size_t j = 0; j < in[0].size(); j++) out[j][i] = in[i][j];
  return out;
}

// mpasoflecsi/eqns/matrix.hh:40-55
Mat matMul(const Mat &A, const Mat &B) {
  Mat C(A.size(), std::vector<double>(B[0].size()));
  for (size_t row = 0; row < C.size(); row++)
    for (size_t col = 0; col < C[0].size(); col++)
      for (size_t inner = 0; inner < B.size(); inner++)
        C[row][col] += A[row][inner] * B[inner][col];
  return C;
}

// mpasoflecsi/eqns/matrix.hh:64-117 (partial-pivoting Gaussian elimination;
// the unqualified abs there resolves to the double overload because
// init_functions.cc includes <stdlib.h>, whose C++ wrapper brings std::abs
// into the global namespace)
void elgs(Mat &a, std::vector<size_t> &idx) {
  size_t n = a.size();
  for (size_t i = 0; i < n; i++) idx[i] = i;
  std::vector<double> c(n);
  double c1;
  for (size_t i = 0; i < n; i++) {
    c1 = 0.;
    for (size_t j = 0; j < n; j++) c1 = std::max(c1, std::abs(a[i][j]));
    c[i] = c1;
  }
  for (size_t j = 0; j < n - 1; j++) {
    size_t k = j;
    double pi1 = 0.0;
    for (size_t i = j; i < n; i++) {
      double pi = std::abs(a[idx[i]][j]) / c[idx[i]];
      if (pi > pi1) {
        pi1 = pi;
        k = i;
      }
    }
    size_t itmp = idx[j];
    idx[j] = idx[k];
    idx[k] = itmp;
    for (size_t i = j + 1; i < n; i++) {
      double pj = a[idx[i]][j] / a[idx[j]][j];
      a[idx[i]][j] = pj;
      for (size_t kk = j + 1; kk < n; kk++) a[idx[i]][kk] = a[idx[i]][kk] - pj * a[idx[j]][kk];
    }
  }
}

// mpasoflecsi/eqns/matrix.hh:137-179 (inverse via elgs)
Mat migs(Mat a) {
  size_t n = a.size();
  Mat b(n, std::vector<double>(n)), x(n, std::vector<double>(n));
  std::vector<size_t> idx(n);
  for (size_t j = 0; j < n; j++)
    for (size_t i = 0; i < n; i++) b[i][j] = 0.;
  for (size_t i = 0; i < n; i++) b[i][i] = 1.;
  elgs(a, idx);
  for (size_t i = 0; i < n - 1; i++)
    for (size_t j = i + 1; j < n; j++)
      for (size_t k = 0; k < n; k++) b[idx[j]][k] = b[idx[j]][k] - a[idx[j]][i] * b[idx[i]][k];
  for (int i = 0; i < int(n); i++) {
    x[n - 1][i] = b[idx[n - 1]][i] / a[idx[n - 1]][n - 1];
    for (int j = int(n) - 2; j >= 0; j--) {
      x[j][i] = b[idx[j]][i];
      for (int k = j + 1; k < int(n); k++) x[j][i] = x[j][i] - a[idx[j]][k] * x[k][i];
      x[j][i] = x[j][i] / a[idx[j]][j];
    }
  }
  return x;
}

// apps/ocean/tasks/init_functions.cc:83-328
void init_2nd_deriv(Mesh &m, double a) {
  const double pi = 3.141592653589793;
  const size_t ll = 25;
  m.derivTwo.assign(size_t(m.nE) * 16 * 2, 0.0);
  auto D2 = [&](int e, int j, int i) -> double & { return m.derivTwo[(size_t(e) * 16 + j) * 2 + i]; };
  const int nCells = m.nC;
  for (int c = 0; c < m.nC; c++) {
    std::vector<int> cell_list{};
    const int nEonC = m.c2e.size(c);
    cell_list.push_back(c);
    for (int j = 0; j < m.c2c.size(c); j++) cell_list.push_back(m.c2c.at(c, j));
    int n = nEonC + 1;
    bool do_the_cell = true;
    for (int i = 1; i < n; i++)
      if (cell_list[i] < 0 or cell_list[i] >= nCells) {
        do_the_cell = false;
        break;
      }
    if (not do_the_cell) continue;

    std::vector<double> xc, yc, zc, theta_v, theta_t, dl_sphere, xp, yp;
    double theta_abs;
    for (int i = 0; i < n; i++) {
      xc.push_back(m.xCell[cell_list[i]] / a);
      yc.push_back(m.yCell[cell_list[i]] / a);
      zc.push_back(m.zCell[cell_list[i]] / a);
    }
    if (zc[0] == 1.0)
      theta_abs = pi / 2.0;
    else
      theta_abs = pi / 2.0 - sphere_angle(xc[0], yc[0], zc[0], xc[1], yc[1], zc[1], 0.0, 0.0, 1.0);
    double length_scale = 1.0;
    for (int i = 0; i < n - 1; i++) {
      int ip2 = i + 2;
      if (ip2 > n - 1) ip2 = 1;
      theta_v.push_back(sphere_angle(xc[0], yc[0], zc[0], xc[i + 1], yc[i + 1], zc[i + 1],
                                     xc[ip2], yc[ip2], zc[ip2]));
      dl_sphere.push_back(a * sphere_arc_length(xc[0], yc[0], zc[0], xc[i + 1], yc[i + 1],
                                                zc[i + 1]) /
                          length_scale);
    }
    theta_t.push_back(theta_abs);
    for (int i = 1; i < n - 1; i++) theta_t.push_back(theta_t[i - 1] + theta_v[i - 1]);
    for (int i = 0; i < n - 1; i++) {
      xp.push_back(std::cos(theta_t[i]) * dl_sphere[i]);
      yp.push_back(std::sin(theta_t[i]) * dl_sphere[i]);
    }

    // mpas_poly_fit_2 (init_functions.cc:206-267)
    size_t na = 6, ma = n;
    static thread_local double amatrix[ll][ll], bmatrix[ll][ll], wmatrix[ll][ll];
    for (size_t i = 0; i < ll; i++)
      for (size_t j = 0; j < ll; j++) amatrix[i][j] = bmatrix[i][j] = wmatrix[i][j] = 0;
    amatrix[0][0] = 1.0;
    wmatrix[0][0] = 1.0;
    for (size_t i = 1; i < ma; i++) {
      amatrix[i][0] = 1.0;
      amatrix[i][1] = xp[i - 1];
      amatrix[i][2] = yp[i - 1];
      amatrix[i][3] = xp[i - 1] * xp[i - 1];
      amatrix[i][4] = xp[i - 1] * yp[i - 1];
      amatrix[i][5] = yp[i - 1] * yp[i - 1];
      wmatrix[i][i] = 1.0;
    }
    Mat aa, ww;
    for (size_t j = 0; j < ma; j++) {
      std::vector<double> atmp;
      for (size_t i = 0; i < na; i++) atmp.push_back(amatrix[j][i]);
      aa.push_back(atmp);
    }
    for (size_t j = 0; j < ma; j++) {
      std::vector<double> wtmp;
      for (size_t i = 0; i < ma; i++) wtmp.push_back(wmatrix[j][i]);
      ww.push_back(wtmp);
    }
    auto at = transpose(aa);
    auto wt = transpose(ww);
    auto h = matMul(wt, ww);
    auto ath = matMul(at, h);
    auto atha = matMul(ath, aa);
    auto atha_inv = migs(atha);
    auto bb = matMul(atha_inv, ath);
    for (size_t j = 0; j < na; j++)
      for (size_t i = 0; i < ma; i++) bmatrix[j][i] = bb[j][i];

    for (int i = 0; i < nEonC; i++) {  // 272-324
      int e1 = m.c2e.at(c, i);
      int va = m.e2v.at(e1, 0), vb = m.e2v.at(e1, 1);
      double xv1 = m.xVertex[va] / a, yv1 = m.yVertex[va] / a, zv1 = m.zVertex[va] / a;
      double xv2 = m.xVertex[vb] / a, yv2 = m.yVertex[vb] / a, zv2 = m.zVertex[vb] / a;
      double xec, yec, zec;
      arc_bisect(xv1, yv1, zv1, xv2, yv2, zv2, xec, yec, zec);
      double thetae_tmp = sphere_angle(xc[0], yc[0], zc[0], xc[i + 1], yc[i + 1], zc[i + 1],
                                       xec, yec, zec);
      thetae_tmp += theta_t[i];
      double cos2t = std::cos(thetae_tmp);
      double sin2t = std::sin(thetae_tmp);
      double costsint = cos2t * sin2t;
      cos2t = cos2t * cos2t;
      sin2t = sin2t * sin2t;
      int col = (c == m.e2c.at(e1, 0)) ? 0 : 1;
      for (int j = 0; j < n; j++)
        D2(e1, j, col) = 2. * cos2t * bmatrix[3][j] + 2. * costsint * bmatrix[4][j] +
                         2. * sin2t * bmatrix[5][j];
    }
  }
}

// apps/ocean/tasks/init_functions.cc:330-447
void init_tracer_adv_coeff(Mesh &m, int nL) {
  m.advCoef.assign(size_t(m.nE) * 16, 0.0);
  m.advCoef_3rd.assign(size_t(m.nE) * 16, 0.0);
  m.advCellsForEdge.assign(size_t(m.nE) * 16, 0);
  m.highOrderAdvMask.assign(size_t(m.nE) * nL, 0);
  m.nAdvCellsForEdge.assign(m.nE, 0);
  auto D2 = [&](int e, int j, int i) { return m.derivTwo[(size_t(e) * 16 + j) * 2 + i]; };
  for (int e = 0; e < m.nE; e++)
    for (int k = 0; k < nL; k++) {
      m.highOrderAdvMask[e * nL + k] = 1;
      for (int j = 0; j < m.e2c.size(e); j++)
        if (m.boundaryCell[m.e2c.at(e, j) * nL + k] == 1) m.highOrderAdvMask[e * nL + k] = 0;
    }
  for (int e = 0; e < m.nE; e++) {
    double *ac = &m.advCoef[size_t(e) * 16];
    double *a3 = &m.advCoef_3rd[size_t(e) * 16];
    int cell1 = m.e2c.at(e, 0), cell2 = m.e2c.at(e, 1);
    std::vector<int> cell_list{cell1, cell2};
    int nEonC1 = m.c2e.size(cell1);
    for (int j = 0; j < m.c2c.size(cell1); j++) {
      int c = m.c2c.at(cell1, j);
      if (std::find(cell_list.begin(), cell_list.end(), c) == cell_list.end()) cell_list.push_back(c);
    }
    int nEonC2 = m.c2e.size(cell2);
    for (int j = 0; j < m.c2c.size(cell2); j++) {
      int c = m.c2c.at(cell2, j);
      if (std::find(cell_list.begin(), cell_list.end(), c) == cell_list.end()) cell_list.push_back(c);
    }
    int nAdv = int(cell_list.size());
    if (nAdv > 16) throw std::runtime_error("more than maxEdges2 advection cells");
    m.nAdvCellsForEdge[e] = nAdv;
    for (int i = 0; i < nAdv; i++) m.advCellsForEdge[size_t(e) * 16 + i] = cell_list[i];
    auto find = [&](int c) {
      return int(std::distance(cell_list.begin(), std::find(cell_list.begin(), cell_list.end(), c)));
    };
    int idx = find(cell1);
    if (idx < nAdv) {
      ac[idx] = ac[idx] + D2(e, 0, 0);
      a3[idx] = a3[idx] + D2(e, 0, 0);
    }
    for (int i = 0; i < nEonC1; i++) {
      idx = find(m.c2c.at(cell1, i));
      if (idx < nAdv) {
        ac[idx] = ac[idx] + D2(e, i + 1, 0);
        a3[idx] = a3[idx] + D2(e, i + 1, 0);
      }
    }
    idx = find(cell2);
    if (idx < nAdv) {
      ac[idx] = ac[idx] + D2(e, 0, 1);
      a3[idx] = a3[idx] - D2(e, 0, 1);
    }
    for (int i = 0; i < nEonC2; i++) {
      idx = find(m.c2c.at(cell2, i));
      if (idx < nAdv) {
        ac[idx] = ac[idx] + D2(e, i + 1, 1);
        a3[idx] = a3[idx] - D2(e, i + 1, 1);
      }
    }
    for (int i = 0; i < nAdv; i++) {
      ac[i] = -(m.dcEdge[e] * m.dcEdge[e]) * ac[i] / 12.0;
      a3[i] = -(m.dcEdge[e] * m.dcEdge[e]) * a3[i] / 12.0;
    }
    idx = find(cell1);
    if (idx < nAdv) ac[idx] = ac[idx] + 0.5;
    idx = find(cell2);
    if (idx < nAdv) ac[idx] = ac[idx] + 0.5;
    for (int i = 0; i < nAdv; i++) {
      ac[i] = m.dvEdge[e] * ac[i];
      a3[i] = m.dvEdge[e] * a3[i];
    }
  }
}

}  // namespace

void migs_dense(int n, const double *a, double *x) {
  Mat m(n, std::vector<double>(n));
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) m[i][j] = a[i * n + j];
  const Mat inv = migs(m);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) x[i * n + j] = inv[i][j];
}

// apps/ocean/initialize.cc:273-287 order: maxLevelEdgeTop, (layer thickness
// edge is a field, set by the caller), meshScalingDel2, edgeSignOnCell,
// 2nd derivatives, advection coefficients.  maxLevelCell and boundaryCell
// come from apps/ocean/tasks/testcase.cc:76-94 (every cell active, no
// boundary) unless the caller set them.
void setup_fct(Mesh &m, int nL, double a) {
  if (m.maxLevelCell.size() != size_t(m.nC)) m.maxLevelCell.assign(m.nC, nL);
  if (m.boundaryCell.size() != size_t(m.nC) * nL) m.boundaryCell.assign(size_t(m.nC) * nL, 0);
  // init_functions.cc:18-27
  m.maxLevelEdgeTop.assign(m.nE, 0);
  for (int e = 0; e < m.nE; e++)
    m.maxLevelEdgeTop[e] = std::max(m.maxLevelCell[m.e2c.at(e, 0)], m.maxLevelCell[m.e2c.at(e, 1)]);
  // init_functions.cc:47-66 (every edge uses cells 0 and 1, lines 60-61)
  double maxDensity = 0.0;
  for (int c = 0; c < m.nC; c++)
    if (m.meshDensity[c] > maxDensity) maxDensity = m.meshDensity[c];
  m.meshScalingDel2.assign(m.nE, 0.0);
  for (int e = 0; e < m.nE; e++) {
    double relativeDensity = (m.meshDensity[0] + m.meshDensity[1]) / (2.0 * maxDensity);
    m.meshScalingDel2[e] = 1.0 / std::pow(relativeDensity, 0.25);
  }
  // init_functions.cc:68-82
  m.edgeSignOnCell.assign(size_t(m.nC) * 8, 0.0);
  for (int c = 0; c < m.nC; c++)
    for (int j = 0; j < m.c2e.size(c); j++)
      m.edgeSignOnCell[c * 8 + j] = (c == m.e2c.at(m.c2e.at(c, j), 0)) ? -1.0 : 1.0;
  init_2nd_deriv(m, a);
  init_tracer_adv_coeff(m, nL);
  m.fct_ready = true;
}

}  // namespace orc
