// One-time host preparation of the mesh: topology flattening and the FCT
// stencil setup.  Compiled with -ffp-contract=off so every a*b+c is two
// rounded operations, the arithmetic of the reference's x86-64 build.
#include "mesh_setup.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <thread>

namespace mpasx {

namespace {

// definition.hh:670-718 (read_connectivity_2D): every non-zero entry of a
// row, rebased from 1 to 0; zeros are padding wherever they appear.
std::string read_rows(const int32_t *buf, int rows, int stride, int n_target,
                      const char *name, Csr &out) {
  out.off.assign(1, 0);
  out.idx.clear();
  out.idx.reserve(size_t(rows) * stride);
  for (int i = 0; i < rows; i++) {
    for (int j = 0; j < stride; j++) {
      int32_t v = buf[size_t(i) * stride + j];
      if (v == 0) continue;
      if (v < 0 || v > n_target)
        return std::string(name) + ": id " + std::to_string(v) + " in row " +
               std::to_string(i + 1) + " is out of range";
      out.idx.push_back(v - 1);
    }
    out.off.push_back(int32_t(out.idx.size()));
  }
  return "";
}

// mesh.cc:108-116 (transpose_cnx): the lists of the target entities are
// filled by walking the sources in ascending id.
void transpose(const Csr &src, int n_to, Csr &out) {
  out.off.assign(size_t(n_to) + 1, 0);
  for (int32_t t : src.idx) out.off[size_t(t) + 1]++;
  for (int i = 0; i < n_to; i++) out.off[i + 1] += out.off[i];
  out.idx.assign(src.idx.size(), 0);
  std::vector<int32_t> fill(out.off.begin(), out.off.end() - 1);
  for (int s = 0; s < src.rows(); s++)
    for (int j = src.off[s]; j < src.off[s + 1]; j++) out.idx[fill[src.idx[j]]++] = s;
}

template <class T>
std::vector<T> copy_or(const T *p, size_t n, T fill) {
  if (p) return std::vector<T>(p, p + n);
  return std::vector<T>(n, fill);
}

}  // namespace

int max_row(const Csr &c) {
  int w = 0;
  for (int i = 0; i < c.rows(); i++) w = std::max(w, c.size(i));
  return w;
}

std::vector<int32_t> to_ell(const Csr &c, int width) {
  const int n = c.rows();
  std::vector<int32_t> out(size_t(width) * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < width; j++)
      out[size_t(j) * n + i] = j < c.size(i) ? c.at(i, j) : i;
  return out;
}

std::string flatten_mesh(const mpas_mesh_desc &d, int nL, HostMesh &m) {
  if (d.nCells <= 0 || d.nEdges <= 0 || d.nVertices <= 0)
    return "mesh sizes must be positive";
  if (d.maxEdges <= 0 || d.maxEdges > 16) return "maxEdges must be in [1,16]";
  if (d.maxEdges2 != 16)
    return "maxEdges2 must be 16 (weightsOnEdge is read with 16 columns, "
           "mpasoflecsi/io/helpers.hh:172-184)";
  if (d.vertexDegree != 3) return "vertexDegree must be 3";
  if (!d.verticesOnCell || !d.edgesOnCell || !d.cellsOnCell || !d.verticesOnEdge ||
      !d.edgesOnEdge)
    return "a connectivity array is NULL";
  if (!d.areaCell || !d.dvEdge || !d.dcEdge || !d.areaTriangle || !d.kiteAreasOnVertex ||
      !d.fVertex || !d.weightsOnEdge)
    return "a mesh field is NULL";
  m.nC = d.nCells;
  m.nE = d.nEdges;
  m.nV = d.nVertices;
  m.nL = nL;
  m.maxEdges = d.maxEdges;
  m.maxEdges2 = d.maxEdges2;
  m.vdeg = d.vertexDegree;
  const int C = m.nC, E = m.nE, V = m.nV;

  // mesh.cc:67-90: c2v, c2e, c2c, e2v, e2e keep the file order; v2c, e2c,
  // v2e are transposes.
  std::string err;
  Csr f_c2v, f_c2e, f_c2c, f_e2v, f_e2e;
  if (!(err = read_rows(d.verticesOnCell, C, d.maxEdges, V, "verticesOnCell", m.c2v)).empty() ||
      !(err = read_rows(d.edgesOnCell, C, d.maxEdges, E, "edgesOnCell", m.c2e)).empty() ||
      !(err = read_rows(d.cellsOnCell, C, d.maxEdges, C, "cellsOnCell", m.c2c)).empty() ||
      !(err = read_rows(d.verticesOnEdge, E, 2, V, "verticesOnEdge", m.e2v)).empty() ||
      !(err = read_rows(d.edgesOnEdge, E, d.maxEdges2, E, "edgesOnEdge", m.e2e)).empty())
    return err;
  transpose(m.c2v, V, m.v2c);
  transpose(m.c2e, E, m.e2c);
  transpose(m.e2v, V, m.v2e);

  // The kernels gather where the reference scatters; that is exact only on
  // a closed mesh whose lists have the sizes the task loops index.
  for (int e = 0; e < E; e++) {
    if (m.e2c.size(e) != 2)
      return "edge " + std::to_string(e) + " belongs to " + std::to_string(m.e2c.size(e)) +
             " cells (a closed spherical mesh has 2)";
    if (m.e2v.size(e) != 2 || m.e2v.at(e, 0) == m.e2v.at(e, 1))
      return "edge " + std::to_string(e) + " needs two distinct verticesOnEdge";
  }
  for (int v = 0; v < V; v++)
    if (m.v2c.size(v) != 3)
      return "vertex " + std::to_string(v) + " belongs to " + std::to_string(m.v2c.size(v)) +
             " cells (vertexDegree is 3)";
  for (int c = 0; c < C; c++) {
    if (m.c2e.size(c) < 3) return "cell " + std::to_string(c) + " has fewer than 3 edges";
    for (int j = 0; j < m.c2e.size(c); j++)
      for (int k = 0; k < j; k++)
        if (m.c2e.at(c, j) == m.c2e.at(c, k))
          return "cell " + std::to_string(c) + " lists an edge twice";
  }

  m.areaCell.assign(d.areaCell, d.areaCell + C);
  m.dvEdge.assign(d.dvEdge, d.dvEdge + E);
  m.dcEdge.assign(d.dcEdge, d.dcEdge + E);
  m.areaTriangle.assign(d.areaTriangle, d.areaTriangle + V);
  m.kite.assign(d.kiteAreasOnVertex, d.kiteAreasOnVertex + size_t(V) * 3);
  m.fVertex.assign(d.fVertex, d.fVertex + V);
  m.weightsOnEdge.assign(d.weightsOnEdge, d.weightsOnEdge + size_t(E) * 16);
  if (d.xCell && d.yCell && d.zCell && d.xVertex && d.yVertex && d.zVertex) {
    m.xCell.assign(d.xCell, d.xCell + C);
    m.yCell.assign(d.yCell, d.yCell + C);
    m.zCell.assign(d.zCell, d.zCell + C);
    m.xVertex.assign(d.xVertex, d.xVertex + V);
    m.yVertex.assign(d.yVertex, d.yVertex + V);
    m.zVertex.assign(d.zVertex, d.zVertex + V);
  }
  m.meshDensity = copy_or(d.meshDensity, size_t(C), 1.0);
  m.maxLevelCell = copy_or(d.maxLevelCell, size_t(C), int32_t(nL));
  m.boundaryCell = copy_or(d.boundaryCell, size_t(C) * nL, int32_t(0));
  for (int c = 0; c < C; c++)
    if (m.maxLevelCell[c] < 0 || m.maxLevelCell[c] > nL)
      return "maxLevelCell must lie in [0, nLevels]";
  return "";
}

// ---------------------------------------------------------------------------
// FCT / del2 setup.  The steps are those of apps/ocean/tasks/init_functions.cc
// with the helpers of mpasoflecsi/eqns/{metrics,matrix}.hh, kept step for
// step (same operands, same association) so the coefficients are the
// reference's to the last bit; only the containers differ.
namespace {

constexpr double kPi = 3.141592653589793;  // mpas_constants::pi

inline double clamp_unit(double x) { return std::max(std::min(x, 1.0), -1.0); }

// metrics.hh:23-73: signed angle at A of the spherical triangle ABC.
double angle_at(double ax, double ay, double az, double bx, double by, double bz,
                double cx, double cy, double cz) {
  const double sa = std::acos(clamp_unit(bx * cx + by * cy + bz * cz));
  const double sb = std::acos(clamp_unit(ax * cx + ay * cy + az * cz));
  const double sc = std::acos(clamp_unit(ax * bx + ay * by + az * bz));
  const double ux = bx - ax, uy = by - ay, uz = bz - az;
  const double vx = cx - ax, vy = cy - ay, vz = cz - az;
  const double nx = (uy * vz) - (uz * vy);
  const double ny = (uz * vx) - (ux * vz);
  const double nz = (ux * vy) - (uy * vx);
  const double s = 0.5 * (sa + sb + sc);
  const double q = std::sqrt(
      std::min(1.0, std::max(0.0, (std::sin(s - sb) * std::sin(s - sc)) /
                                      (std::sin(sb) * std::sin(sc)))));
  const double half = std::asin(clamp_unit(q));
  return (nx * ax + ny * ay + nz * az >= 0.0) ? 2.0 * half : -2.0 * half;
}

// metrics.hh:75-91: great-circle length from the chord.
double arc_from_chord(double ax, double ay, double az, double bx, double by, double bz) {
  const double dx = bx - ax, dy = by - ay, dz = bz - az;
  const double r = std::sqrt(ax * ax + ay * ay + az * az);
  const double ch = std::sqrt(dx * dx + dy * dy + dz * dz);
  return r * 2. * std::asin(ch / (2. * r));
}

// metrics.hh:93-111: midpoint of the arc AB, on the sphere of |A|.
bool arc_midpoint(double ax, double ay, double az, double bx, double by, double bz,
                  double &cx, double &cy, double &cz) {
  const double r = std::sqrt(ax * ax + ay * ay + az * az);
  cx = 0.5 * (ax + bx);
  cy = 0.5 * (ay + by);
  cz = 0.5 * (az + bz);
  if (!(cx != 0. && cy != 0. && cz != 0.)) return false;  // the reference's flog_assert
  const double dd = std::sqrt(cx * cx + cy * cy + cz * cz);
  cx = r * cx / dd;
  cy = r * cy / dd;
  cz = r * cz / dd;
  return true;
}

constexpr int kMaxFit = 17;  // 1 + maxEdges (16)

// Dense row-major matrix with the reference's zero-initialised accumulate
// (matrix.hh:40-55: C[r][c] += A[r][i] * B[i][c] from 0).
struct Mat {
  int r = 0, c = 0;
  double v[kMaxFit][kMaxFit];
  void zero(int rows, int cols) {
    r = rows;
    c = cols;
    for (int i = 0; i < rows; i++)
      for (int j = 0; j < cols; j++) v[i][j] = 0.0;
  }
};

void mat_t(const Mat &a, Mat &out) {
  out.zero(a.c, a.r);
  for (int i = 0; i < a.r; i++)
    for (int j = 0; j < a.c; j++) out.v[j][i] = a.v[i][j];
}

void mat_mul(const Mat &a, const Mat &b, Mat &out) {
  out.zero(a.r, b.c);
  for (int i = 0; i < out.r; i++)
    for (int j = 0; j < out.c; j++)
      for (int k = 0; k < b.r; k++) out.v[i][j] += a.v[i][k] * b.v[k][j];
}

// matrix.hh:64-117 and 137-179: scaled partial pivoting elimination
// (T. Pang, An Introduction to Computational Physics, 1997, programs 4.3/4.4)
// and the inverse built from it.
void invert(Mat a, Mat &x) {
  const int n = a.r;
  int idx[kMaxFit];
  double scale[kMaxFit];
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 0; i < n; i++) {
    double mx = 0.;
    for (int j = 0; j < n; j++) mx = std::max(mx, std::fabs(a.v[i][j]));
    scale[i] = mx;
  }
  for (int j = 0; j < n - 1; j++) {
    int piv = j;
    double best = 0.0;
    for (int i = j; i < n; i++) {
      const double ratio = std::fabs(a.v[idx[i]][j]) / scale[idx[i]];
      if (ratio > best) {
        best = ratio;
        piv = i;
      }
    }
    std::swap(idx[j], idx[piv]);
    for (int i = j + 1; i < n; i++) {
      const double f = a.v[idx[i]][j] / a.v[idx[j]][j];
      a.v[idx[i]][j] = f;
      for (int k = j + 1; k < n; k++) a.v[idx[i]][k] = a.v[idx[i]][k] - f * a.v[idx[j]][k];
    }
  }
  Mat b;
  b.zero(n, n);
  for (int i = 0; i < n; i++) b.v[i][i] = 1.;
  for (int i = 0; i < n - 1; i++)
    for (int j = i + 1; j < n; j++)
      for (int k = 0; k < n; k++) b.v[idx[j]][k] = b.v[idx[j]][k] - a.v[idx[j]][i] * b.v[idx[i]][k];
  x.zero(n, n);
  for (int i = 0; i < n; i++) {
    x.v[n - 1][i] = b.v[idx[n - 1]][i] / a.v[idx[n - 1]][n - 1];
    for (int j = n - 2; j >= 0; j--) {
      x.v[j][i] = b.v[idx[j]][i];
      for (int k = j + 1; k < n; k++) x.v[j][i] = x.v[j][i] - a.v[idx[j]][k] * x.v[k][i];
      x.v[j][i] = x.v[j][i] / a.v[idx[j]][j];
    }
  }
}

// init_functions.cc:83-328 for one cell: quadratic least-squares fit on the
// tangent plane, then the second derivative normal to each edge.
std::string second_derivs_of_cell(HostMesh &m, int c, double a) {
  const int nEonC = m.c2e.size(c);
  const int n = nEonC + 1;
  if (m.c2c.size(c) < nEonC)
    return "cell " + std::to_string(c) + " has fewer cellsOnCell than edgesOnCell";
  int list[kMaxFit];
  list[0] = c;
  for (int i = 1; i < n; i++) list[i] = m.c2c.at(c, i - 1);
  double xc[kMaxFit], yc[kMaxFit], zc[kMaxFit];
  for (int i = 0; i < n; i++) {
    xc[i] = m.xCell[list[i]] / a;
    yc[i] = m.yCell[list[i]] / a;
    zc[i] = m.zCell[list[i]] / a;
  }
  double theta_abs;
  if (zc[0] == 1.0)
    theta_abs = kPi / 2.0;
  else
    theta_abs = kPi / 2.0 - angle_at(xc[0], yc[0], zc[0], xc[1], yc[1], zc[1], 0.0, 0.0, 1.0);
  double theta_v[kMaxFit], dl[kMaxFit], theta_t[kMaxFit], xp[kMaxFit], yp[kMaxFit];
  for (int i = 0; i < n - 1; i++) {
    int ip2 = i + 2;
    if (ip2 > n - 1) ip2 = 1;
    theta_v[i] = angle_at(xc[0], yc[0], zc[0], xc[i + 1], yc[i + 1], zc[i + 1], xc[ip2],
                          yc[ip2], zc[ip2]);
    dl[i] = a * arc_from_chord(xc[0], yc[0], zc[0], xc[i + 1], yc[i + 1], zc[i + 1]) / 1.0;
  }
  theta_t[0] = theta_abs;
  for (int i = 1; i < n - 1; i++) theta_t[i] = theta_t[i - 1] + theta_v[i - 1];
  for (int i = 0; i < n - 1; i++) {
    xp[i] = std::cos(theta_t[i]) * dl[i];
    yp[i] = std::sin(theta_t[i]) * dl[i];
  }

  // mpas_poly_fit_2 (init_functions.cc:206-267): B = (A^T W^T W A)^-1 A^T W^T W
  const int na = 6, ma = n;
  Mat A, W;
  A.zero(ma, na);
  W.zero(ma, ma);
  A.v[0][0] = 1.0;
  W.v[0][0] = 1.0;
  for (int i = 1; i < ma; i++) {
    A.v[i][0] = 1.0;
    A.v[i][1] = xp[i - 1];
    A.v[i][2] = yp[i - 1];
    A.v[i][3] = xp[i - 1] * xp[i - 1];
    A.v[i][4] = xp[i - 1] * yp[i - 1];
    A.v[i][5] = yp[i - 1] * yp[i - 1];
    W.v[i][i] = 1.0;
  }
  Mat At, Wt, H, AtH, AtHA, AtHAinv, B;
  mat_t(A, At);
  mat_t(W, Wt);
  mat_mul(Wt, W, H);
  mat_mul(At, H, AtH);
  mat_mul(AtH, A, AtHA);
  invert(AtHA, AtHAinv);
  mat_mul(AtHAinv, AtH, B);

  for (int i = 0; i < nEonC; i++) {
    const int e1 = m.c2e.at(c, i);
    const int va = m.e2v.at(e1, 0), vb = m.e2v.at(e1, 1);
    const double xv1 = m.xVertex[va] / a, yv1 = m.yVertex[va] / a, zv1 = m.zVertex[va] / a;
    const double xv2 = m.xVertex[vb] / a, yv2 = m.yVertex[vb] / a, zv2 = m.zVertex[vb] / a;
    double xec, yec, zec;
    if (!arc_midpoint(xv1, yv1, zv1, xv2, yv2, zv2, xec, yec, zec))
      return "arc_bisect: the midpoint of edge " + std::to_string(e1) +
             " has a zero coordinate (metrics.hh:104-105 asserts)";
    double thetae = angle_at(xc[0], yc[0], zc[0], xc[i + 1], yc[i + 1], zc[i + 1], xec, yec, zec);
    thetae += theta_t[i];
    double cos2t = std::cos(thetae);
    double sin2t = std::sin(thetae);
    const double cs = cos2t * sin2t;
    cos2t = cos2t * cos2t;
    sin2t = sin2t * sin2t;
    const int col = (c == m.e2c.at(e1, 0)) ? 0 : 1;
    for (int j = 0; j < n; j++)
      m.derivTwo[(size_t(e1) * 16 + j) * 2 + col] =
          2. * cos2t * B.v[3][j] + 2. * cs * B.v[4][j] + 2. * sin2t * B.v[5][j];
  }
  return "";
}

// init_functions.cc:330-447 for one edge.
std::string adv_coefficients_of_edge(HostMesh &m, int e) {
  double *ac = &m.advCoef[size_t(e) * 16];
  double *a3 = &m.advCoef_3rd[size_t(e) * 16];
  int *cells = &m.advCellsForEdge[size_t(e) * 16];
  const int c1 = m.e2c.at(e, 0), c2 = m.e2c.at(e, 1);
  int list[32];
  int n = 0;
  auto find = [&](int x) {
    for (int i = 0; i < n; i++)
      if (list[i] == x) return i;
    return n;
  };
  list[n++] = c1;
  list[n++] = c2;
  for (int side = 0; side < 2; side++) {
    const int cc = side ? c2 : c1;
    for (int j = 0; j < m.c2c.size(cc); j++) {
      const int x = m.c2c.at(cc, j);
      if (find(x) == n) {
        if (n == 32) return "too many advection cells";
        list[n++] = x;
      }
    }
  }
  if (n > 16)
    return "edge " + std::to_string(e) + " needs " + std::to_string(n) +
           " advection cells (nAdvectionCells = maxEdges2 = 16)";
  m.nAdvCellsForEdge[e] = n;
  for (int i = 0; i < n; i++) cells[i] = list[i];
  auto D2 = [&](int j, int col) { return m.derivTwo[(size_t(e) * 16 + j) * 2 + col]; };
  int k = find(c1);
  if (k < n) {
    ac[k] = ac[k] + D2(0, 0);
    a3[k] = a3[k] + D2(0, 0);
  }
  const int n1 = m.c2e.size(c1), n2 = m.c2e.size(c2);
  for (int i = 0; i < n1; i++) {
    k = find(m.c2c.at(c1, i));
    if (k < n) {
      ac[k] = ac[k] + D2(i + 1, 0);
      a3[k] = a3[k] + D2(i + 1, 0);
    }
  }
  k = find(c2);
  if (k < n) {
    ac[k] = ac[k] + D2(0, 1);
    a3[k] = a3[k] - D2(0, 1);
  }
  for (int i = 0; i < n2; i++) {
    k = find(m.c2c.at(c2, i));
    if (k < n) {
      ac[k] = ac[k] + D2(i + 1, 1);
      a3[k] = a3[k] - D2(i + 1, 1);
    }
  }
  const double dc = m.dcEdge[e], dv = m.dvEdge[e];
  for (int i = 0; i < n; i++) {
    ac[i] = -(dc * dc) * ac[i] / 12.0;
    a3[i] = -(dc * dc) * a3[i] / 12.0;
  }
  k = find(c1);
  if (k < n) ac[k] = ac[k] + 0.5;
  k = find(c2);
  if (k < n) ac[k] = ac[k] + 0.5;
  for (int i = 0; i < n; i++) {
    ac[i] = dv * ac[i];
    a3[i] = dv * a3[i];
  }
  return "";
}

// Runs f(i) for i in [0, n) on the host cores; returns the first error.
template <class F>
std::string parallel_for(int n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::string> errs(nt);
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; t++)
    pool.emplace_back([&, t] {
      const int lo = int(int64_t(n) * t / nt), hi = int(int64_t(n) * (t + 1) / nt);
      for (int i = lo; i < hi && errs[t].empty(); i++) errs[t] = f(i);
    });
  for (auto &th : pool) th.join();
  for (auto &e : errs)
    if (!e.empty()) return e;
  return "";
}

}  // namespace

std::string setup_tracer_operators(HostMesh &m, double a) {
  const int C = m.nC, E = m.nE, nL = m.nL;
  if (m.xCell.empty()) return "the FCT setup needs xCell/yCell/zCell and xVertex/yVertex/zVertex";
  for (int c = 0; c < C; c++) {
    if (m.c2e.size(c) > 8) return "edgeSignOnCell holds maxEdges = 8 entries per cell";
    if (m.c2c.size(c) != m.c2e.size(c))
      return "cell " + std::to_string(c) + " lists different numbers of cellsOnCell and edgesOnCell";
  }
  // init_functions.cc:18-27
  m.maxLevelEdgeTop.assign(E, 0);
  for (int e = 0; e < E; e++)
    m.maxLevelEdgeTop[e] =
        std::max(m.maxLevelCell[m.e2c.at(e, 0)], m.maxLevelCell[m.e2c.at(e, 1)]);
  // init_functions.cc:47-66.  Lines 60-61 take cells 0 and 1 of the whole
  // mesh for every edge, so the scaling is one number for all edges.
  double maxDensity = 0.0;
  for (int c = 0; c < C; c++)
    if (m.meshDensity[c] > maxDensity) maxDensity = m.meshDensity[c];
  const double rel = (m.meshDensity[0] + m.meshDensity[std::min(1, C - 1)]) / (2.0 * maxDensity);
  m.meshScalingDel2.assign(E, 1.0 / std::pow(rel, 0.25));
  // init_functions.cc:68-82
  m.edgeSignOnCell.assign(size_t(C) * 8, 0.0);
  for (int c = 0; c < C; c++)
    for (int j = 0; j < m.c2e.size(c); j++)
      m.edgeSignOnCell[size_t(c) * 8 + j] = (c == m.e2c.at(m.c2e.at(c, j), 0)) ? -1.0 : 1.0;
  // init_functions.cc:83-328
  m.derivTwo.assign(size_t(E) * 32, 0.0);
  for (int c = 0; c < C; c++)
    for (int i = 0; i < m.c2c.size(c) && i < m.c2e.size(c); i++)
      if (m.c2c.at(c, i) < 0 || m.c2c.at(c, i) >= C)
        return "cellsOnCell holds an id outside the mesh";
  std::string err = parallel_for(C, [&](int c) { return second_derivs_of_cell(m, c, a); });
  if (!err.empty()) return err;
  // init_functions.cc:330-447
  m.advCoef.assign(size_t(E) * 16, 0.0);
  m.advCoef_3rd.assign(size_t(E) * 16, 0.0);
  m.advCellsForEdge.assign(size_t(E) * 16, 0);
  m.nAdvCellsForEdge.assign(E, 0);
  m.highOrderAdvMask.assign(size_t(E) * nL, 0);
  for (int e = 0; e < E; e++)
    for (int k = 0; k < nL; k++) {
      int mask = 1;
      for (int j = 0; j < 2; j++)
        if (m.boundaryCell[size_t(m.e2c.at(e, j)) * nL + k] == 1) mask = 0;
      m.highOrderAdvMask[size_t(e) * nL + k] = mask;
    }
  err = parallel_for(E, [&](int e) { return adv_coefficients_of_edge(m, e); });
  if (!err.empty()) return err;
  // The reference normalises positions by its constant a; with a radius
  // that does not match the mesh, neighbour dot products exceed 1, every
  // clamped angle is 0 and the fit divides 0 by 0.  Refuse such a stencil.
  for (size_t i = 0; i < m.advCoef.size(); i++)
    if (!std::isfinite(m.advCoef[i]) || !std::isfinite(m.advCoef_3rd[i]))
      return "the FCT stencil of edge " + std::to_string(i / 16) +
             " is not finite: sphere_radius must be the mesh's radius";
  m.fct = true;
  return "";
}

namespace {

// Cells in reverse Cuthill-McKee order: breadth-first from a peripheral
// cell (the last one reached by a sweep from cell 0), neighbours by
// ascending degree then id, the whole sequence reversed.
std::vector<int32_t> rcm_order(const HostMesh &m) {
  const int C = m.nC;
  auto bfs = [&](int start, std::vector<int32_t> &seq) {
    std::vector<char> seen(C, 0);
    seq.clear();
    for (int root = start, done = 0; done < C; root = 0) {
      while (root < C && seen[root]) root++;
      if (root >= C) break;
      seen[root] = 1;
      size_t head = seq.size();
      seq.push_back(root);
      while (head < seq.size()) {
        const int c = seq[head++];
        std::vector<int> nb;
        for (int j = 0; j < m.c2c.size(c); j++)
          if (!seen[m.c2c.at(c, j)]) nb.push_back(m.c2c.at(c, j));
        std::sort(nb.begin(), nb.end(), [&](int a, int b) {
          return m.c2c.size(a) != m.c2c.size(b) ? m.c2c.size(a) < m.c2c.size(b) : a < b;
        });
        for (int x : nb)
          if (!seen[x]) seen[x] = 1, seq.push_back(x);
      }
      done = int(seq.size());
    }
  };
  std::vector<int32_t> seq;
  bfs(0, seq);
  bfs(seq.back(), seq);
  std::reverse(seq.begin(), seq.end());
  return seq;
}

// Cells along a 3-D Morton (Z-order) curve of their unit positions.
std::vector<int32_t> morton_order(const HostMesh &m) {
  const int C = m.nC;
  auto spread = [](uint64_t x) {  // 21 bits -> every third bit of 63
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
  };
  std::vector<std::pair<uint64_t, int32_t>> key(C);
  for (int c = 0; c < C; c++) {
    const double r = std::sqrt(m.xCell[c] * m.xCell[c] + m.yCell[c] * m.yCell[c] +
                               m.zCell[c] * m.zCell[c]);
    auto q = [&](double v) { return uint64_t(std::min(2097151.0, (v / r + 1.0) * 1048575.5)); };
    key[c] = {spread(q(m.xCell[c])) | spread(q(m.yCell[c])) << 1 | spread(q(m.zCell[c])) << 2, c};
  }
  std::sort(key.begin(), key.end());
  std::vector<int32_t> seq(C);
  for (int i = 0; i < C; i++) seq[i] = key[i].second;
  return seq;
}

// Rows of `in` in the order `of` (new -> old), entries renamed by `map`
// (old -> new), each row's entry order kept.
Csr remap_rows(const Csr &in, const std::vector<int32_t> &of, const std::vector<int32_t> &map) {
  Csr out;
  out.off.assign(1, 0);
  out.idx.reserve(in.idx.size());
  for (int32_t o : of) {
    for (int j = 0; j < in.size(o); j++) out.idx.push_back(map[in.at(o, j)]);
    out.off.push_back(int32_t(out.idx.size()));
  }
  return out;
}

template <class T>
std::vector<T> move_rows(const std::vector<T> &a, const std::vector<int32_t> &of, size_t w) {
  if (a.empty()) return a;
  std::vector<T> out(a.size());
  for (size_t i = 0; i < of.size(); i++)
    std::copy(a.begin() + size_t(of[i]) * w, a.begin() + size_t(of[i] + 1) * w,
              out.begin() + i * w);
  return out;
}

}  // namespace

std::string reorder_host_mesh(const HostMesh &m, int order, HostMesh &o) {
  const int C = m.nC, E = m.nE, V = m.nV;
  std::vector<int32_t> cof;
  if (order == kOrderGiven) {
    cof.resize(C);
    for (int c = 0; c < C; c++) cof[c] = c;
  } else if (order == kOrderRcm) {
    cof = rcm_order(m);
  } else if (order == kOrderMorton) {
    if (m.xCell.size() != size_t(C)) return "the Morton order needs the cell positions";
    cof = morton_order(m);
  } else {
    return "unknown cell order";
  }
  if (int(cof.size()) != C) return "cell order does not cover the mesh";
  std::vector<int32_t> cnew(C);
  for (int i = 0; i < C; i++) cnew[cof[i]] = i;
  // edges by their (lower, higher) new cell pair, vertices by their sorted
  // new cell triple (ties by old id)
  std::vector<int32_t> eof(E), vof(V);
  {
    std::vector<std::pair<std::pair<int, int>, int>> k(E);
    for (int e = 0; e < E; e++) {
      const int a = cnew[m.e2c.at(e, 0)], b = cnew[m.e2c.at(e, 1)];
      k[e] = {{std::min(a, b), std::max(a, b)}, e};
    }
    std::sort(k.begin(), k.end());
    for (int i = 0; i < E; i++) eof[i] = k[i].second;
    std::vector<std::pair<std::array<int, 3>, int>> kv(V);
    for (int v = 0; v < V; v++) {
      std::array<int, 3> t{};
      for (int i = 0; i < 3 && i < m.v2c.size(v); i++) t[i] = cnew[m.v2c.at(v, i)];
      std::sort(t.begin(), t.end());
      kv[v] = {t, v};
    }
    std::sort(kv.begin(), kv.end());
    for (int i = 0; i < V; i++) vof[i] = kv[i].second;
  }
  std::vector<int32_t> enew(E), vnew(V);
  for (int i = 0; i < E; i++) enew[eof[i]] = i;
  for (int i = 0; i < V; i++) vnew[vof[i]] = i;

  o = HostMesh();
  o.nC = C, o.nE = E, o.nV = V, o.nL = m.nL;
  o.maxEdges = m.maxEdges, o.maxEdges2 = m.maxEdges2, o.vdeg = m.vdeg;
  o.c2v = remap_rows(m.c2v, cof, vnew);
  o.c2e = remap_rows(m.c2e, cof, enew);
  o.c2c = remap_rows(m.c2c, cof, cnew);
  o.e2v = remap_rows(m.e2v, eof, vnew);
  o.e2e = remap_rows(m.e2e, eof, enew);
  o.e2c = remap_rows(m.e2c, eof, cnew);
  o.v2c = remap_rows(m.v2c, vof, cnew);
  o.v2e = remap_rows(m.v2e, vof, enew);
  o.areaCell = move_rows(m.areaCell, cof, 1);
  o.xCell = move_rows(m.xCell, cof, 1);
  o.yCell = move_rows(m.yCell, cof, 1);
  o.zCell = move_rows(m.zCell, cof, 1);
  o.meshDensity = move_rows(m.meshDensity, cof, 1);
  o.maxLevelCell = move_rows(m.maxLevelCell, cof, 1);
  o.boundaryCell = move_rows(m.boundaryCell, cof, size_t(m.nL));
  o.edgeSignOnCell = move_rows(m.edgeSignOnCell, cof, 8);
  o.dvEdge = move_rows(m.dvEdge, eof, 1);
  o.dcEdge = move_rows(m.dcEdge, eof, 1);
  o.weightsOnEdge = move_rows(m.weightsOnEdge, eof, 16);
  o.maxLevelEdgeTop = move_rows(m.maxLevelEdgeTop, eof, 1);
  o.nAdvCellsForEdge = move_rows(m.nAdvCellsForEdge, eof, 1);
  o.advCellsForEdge = move_rows(m.advCellsForEdge, eof, 16);
  for (size_t e = 0; e < o.nAdvCellsForEdge.size(); e++)
    for (int i = 0; i < o.nAdvCellsForEdge[e]; i++)
      o.advCellsForEdge[e * 16 + i] = cnew[o.advCellsForEdge[e * 16 + i]];
  o.highOrderAdvMask = move_rows(m.highOrderAdvMask, eof, size_t(m.nL));
  o.meshScalingDel2 = move_rows(m.meshScalingDel2, eof, 1);
  o.derivTwo = move_rows(m.derivTwo, eof, 32);
  o.advCoef = move_rows(m.advCoef, eof, 16);
  o.advCoef_3rd = move_rows(m.advCoef_3rd, eof, 16);
  o.areaTriangle = move_rows(m.areaTriangle, vof, 1);
  o.kite = move_rows(m.kite, vof, 3);
  o.fVertex = move_rows(m.fVertex, vof, 1);
  o.xVertex = move_rows(m.xVertex, vof, 1);
  o.yVertex = move_rows(m.yVertex, vof, 1);
  o.zVertex = move_rows(m.zVertex, vof, 1);
  o.fct = m.fct;
  auto compose = [](const std::vector<int32_t> &prev, const std::vector<int32_t> &of) {
    std::vector<int32_t> r(of.size());
    for (size_t i = 0; i < of.size(); i++) r[i] = prev.empty() ? of[i] : prev[of[i]];
    return r;
  };
  o.cellOf = compose(m.cellOf, cof);
  o.edgeOf = compose(m.edgeOf, eof);
  o.vertexOf = compose(m.vertexOf, vof);
  return "";
}

}  // namespace mpasx
