// Icosahedral spherical Voronoi (TRiSK) mesh generator.  See meshgen.h.
//
// Construction:
//  1. icosahedron (a vertex at each pole, then a fixed generic rotation),
//     bisected `level` times, every new point projected to the unit sphere
//     (these are the cell generators, the triangles are the Voronoi vertices);
//  2. `lloyd_iters` Lloyd iterations towards a centroidal Voronoi
//     tessellation with the triangulation held fixed (checked to stay
//     Delaunay and well centred);
//  3. numbering: cells along a 3-D Hilbert curve, edges sorted by
//     (cellsOnEdge(1), cellsOnEdge(2)), vertices by their sorted cell triple;
//  4. MPAS conventions: cellsOnEdge(1) < cellsOnEdge(2) and the edge normal
//     points from cellsOnEdge(1) to cellsOnEdge(2); verticesOnEdge is ordered
//     so that the tangent k x n points from verticesOnEdge(1) to (2)
//     (the orientation apps/sw/tasks/solve.cc:47-55 assumes for the
//     circulation); cellsOnCell/edgesOnCell/verticesOnCell are
//     counter-clockwise, verticesOnCell(j) sitting between edgesOnCell(j) and
//     edgesOnCell(j+1); cellsOnVertex is written in ascending cell id, so the
//     file order of kiteAreasOnVertex coincides with the transposed order the
//     reference pairs it with (apps/sw/tasks/solve.cc:86).
//  5. TRiSK weights (Thuburn et al. 2009, Ringler et al. 2010):
//       w(e,e') = n(e,i) n(e',i) (1/2 - sum_{v passed} kite(v,i)/area(i))
//                 * dvEdge(e') / dcEdge(e)
//     for e' on cell i of edge e, walking counter-clockwise from e, cell 1
//     first; n(e,i) = +1 when e's normal leaves cell i.  With these weights
//     sum_j w(e,j) u(e_j) reconstructs the tangential velocity along k x n,
//     the sign apps/sw/tasks/solve.cc:215-231 needs for q.
#include "meshgen.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 unit(V3 a) { return (1.0 / norm(a)) * a; }

// great-circle angle between unit vectors
inline double arc(V3 a, V3 b) { return std::atan2(norm(cross(a, b)), dot(a, b)); }

// signed area of the spherical triangle abc on the unit sphere
// (Eriksson 1990: tan(E/2) = a.(b x c) / (1 + a.b + b.c + c.a))
inline double tri_area(V3 a, V3 b, V3 c) {
  double det = dot(a, cross(b, c));
  double den = 1.0 + dot(a, b) + dot(b, c) + dot(c, a);
  return 2.0 * std::atan2(det, den);
}

thread_local std::string g_err;

// Skilling's transform: 3 axes of `bits` bits -> Hilbert index.
uint64_t hilbert3(uint32_t X[3], int bits) {
  const int n = 3;
  uint32_t M = 1u << (bits - 1), P, Q, t;
  for (Q = M; Q > 1; Q >>= 1) {
    P = Q - 1;
    for (int i = 0; i < n; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < n; i++) X[i] ^= X[i - 1];
  t = 0;
  for (Q = M; Q > 1; Q >>= 1)
    if (X[n - 1] & Q) t ^= Q - 1;
  for (int i = 0; i < n; i++) X[i] ^= t;
  uint64_t key = 0;
  for (int b = bits - 1; b >= 0; b--)
    for (int i = 0; i < n; i++) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}

uint64_t hilbert_key(V3 p) {
  const int bits = 21;
  const double scale = double((1u << bits) - 1);
  uint32_t X[3];
  double c[3] = {p.x, p.y, p.z};
  for (int i = 0; i < 3; i++) {
    double u = 0.5 * (c[i] + 1.0);
    u = std::min(1.0, std::max(0.0, u));
    X[i] = uint32_t(u * scale);
  }
  return hilbert3(X, bits);
}

}  // namespace

struct mg_mesh {
  int64_t nCells = 0, nEdges = 0, nVertices = 0;
  std::map<std::string, std::vector<int32_t>> ints;
  std::map<std::string, std::vector<double>> dbls;
};

namespace {

constexpr int kMaxEdges = 8;
constexpr int kMaxEdges2 = 16;

struct Tri {
  int v[3];
};

void build(mg_mesh &M, int level, double radius, int lloyd_iters, int order,
           double omega, int rotate) {
  if (level < 0 || level > 10) throw std::runtime_error("level must be in [0,10]");
  if (!(radius > 0)) throw std::runtime_error("radius must be positive");

  // ---- 1. icosahedron with poles -------------------------------------
  std::vector<V3> P;
  std::vector<Tri> T;
  const double pi = 3.141592653589793;
  const double lat0 = std::atan(0.5);
  P.push_back({0, 0, 1});
  for (int k = 0; k < 5; k++) {
    double lon = 2 * pi * k / 5;
    P.push_back({std::cos(lat0) * std::cos(lon), std::cos(lat0) * std::sin(lon),
                 std::sin(lat0)});
  }
  for (int k = 0; k < 5; k++) {
    double lon = 2 * pi * k / 5 + pi / 5;
    P.push_back({std::cos(lat0) * std::cos(lon), std::cos(lat0) * std::sin(lon),
                 -std::sin(lat0)});
  }
  P.push_back({0, 0, -1});
  // A fixed generic rotation keeps every generator, Voronoi vertex and edge
  // midpoint off the coordinate planes: the reference's FCT setup asserts
  // that no arc midpoint has a zero coordinate (mpasoflecsi/eqns/metrics.hh:
  // 104-105), which the symmetric icosahedron violates from level 6 on.
  if (rotate) {
    const double ax = 0.3117, ay = 0.1731, az = 0.0953;  // radians
    const double cx = std::cos(ax), sx = std::sin(ax), cy = std::cos(ay), sy = std::sin(ay);
    const double cz = std::cos(az), sz = std::sin(az);
    for (auto &p : P) {
      V3 q = {p.x, cx * p.y - sx * p.z, sx * p.y + cx * p.z};
      q = {cy * q.x + sy * q.z, q.y, -sy * q.x + cy * q.z};
      p = unit({cz * q.x - sz * q.y, sz * q.x + cz * q.y, q.z});
    }
  }
  auto U = [](int k) { return 1 + (k % 5); };
  auto L = [](int k) { return 6 + (k % 5); };
  for (int k = 0; k < 5; k++) {
    T.push_back({{0, U(k), U(k + 1)}});
    T.push_back({{U(k), L(k), U(k + 1)}});
    T.push_back({{U(k + 1), L(k), L(k + 1)}});
    T.push_back({{11, L(k + 1), L(k)}});
  }
  for (auto &t : T) {  // enforce counter-clockwise seen from outside
    V3 a = P[t.v[0]], b = P[t.v[1]], c = P[t.v[2]];
    if (dot(cross(b - a, c - a), a + b + c) < 0) std::swap(t.v[1], t.v[2]);
  }

  // ---- bisection -------------------------------------------------------
  for (int l = 0; l < level; l++) {
    std::vector<std::vector<std::pair<int, int>>> mid(P.size());
    auto midpoint = [&](int a, int b) {
      int lo = std::min(a, b), hi = std::max(a, b);
      for (auto &pr : mid[lo])
        if (pr.first == hi) return pr.second;
      int id = int(P.size());
      P.push_back(unit(P[lo] + P[hi]));
      mid[lo].push_back({hi, id});
      return id;
    };
    std::vector<Tri> T2;
    T2.reserve(T.size() * 4);
    for (auto &t : T) {
      int a = t.v[0], b = t.v[1], c = t.v[2];
      int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
      T2.push_back({{a, ab, ca}});
      T2.push_back({{ab, b, bc}});
      T2.push_back({{ca, bc, c}});
      T2.push_back({{ab, bc, ca}});
    }
    T.swap(T2);
  }
  const int nC = int(P.size());
  const int nV = int(T.size());
  if (int64_t(nC) != 10 * (int64_t(1) << (2 * level)) + 2)
    throw std::runtime_error("unexpected cell count");

  // ---- rings: counter-clockwise neighbours and triangles of each generator
  std::vector<int> inc_off(nC + 1, 0), inc;
  for (auto &t : T)
    for (int k = 0; k < 3; k++) inc_off[t.v[k] + 1]++;
  for (int i = 0; i < nC; i++) inc_off[i + 1] += inc_off[i];
  inc.resize(inc_off[nC]);
  {
    std::vector<int> fill(inc_off.begin(), inc_off.end() - 1);
    for (int t = 0; t < nV; t++)
      for (int k = 0; k < 3; k++) inc[fill[T[t].v[k]]++] = t;
  }
  // ring_nb[i][m] = j_m, ring_tri[i][m] = triangle (i, j_m, j_{m+1})
  std::vector<int> ring_off(nC + 1, 0), ring_nb, ring_tri;
  for (int i = 0; i < nC; i++) ring_off[i + 1] = ring_off[i] + (inc_off[i + 1] - inc_off[i]);
  ring_nb.resize(ring_off[nC]);
  ring_tri.resize(ring_off[nC]);
  auto after = [&](int t, int i, int s) {  // vertex s positions after i in tri t
    for (int k = 0; k < 3; k++)
      if (T[t].v[k] == i) return T[t].v[(k + s) % 3];
    return -1;
  };
  for (int i = 0; i < nC; i++) {
    int deg = inc_off[i + 1] - inc_off[i];
    if (deg > kMaxEdges) throw std::runtime_error("cell degree exceeds maxEdges");
    int t = inc[inc_off[i]];
    for (int m = 0; m < deg; m++) {
      ring_nb[ring_off[i] + m] = after(t, i, 1);
      ring_tri[ring_off[i] + m] = t;
      int next = after(t, i, 2);
      int found = -1;
      for (int q = inc_off[i]; q < inc_off[i + 1]; q++)
        if (after(inc[q], i, 1) == next) found = inc[q];
      if (found < 0) throw std::runtime_error("broken triangle ring");
      t = found;
    }
    if (t != inc[inc_off[i]]) throw std::runtime_error("triangle ring does not close");
  }

  // ---- 2. Lloyd iterations ----------------------------------------------
  std::vector<V3> cc(nV);
  auto circumcentres = [&]() {
    for (int t = 0; t < nV; t++) {
      V3 a = P[T[t].v[0]], b = P[T[t].v[1]], c = P[T[t].v[2]];
      cc[t] = unit(cross(b - a, c - a));
    }
  };
  circumcentres();
  for (int it = 0; it < lloyd_iters; it++) {
    std::vector<V3> Pn(nC);
    for (int i = 0; i < nC; i++) {
      int deg = ring_off[i + 1] - ring_off[i];
      V3 s = {0, 0, 0};
      for (int m = 0; m < deg; m++) {
        V3 a = cc[ring_tri[ring_off[i] + m]];
        V3 b = cc[ring_tri[ring_off[i] + (m + 1) % deg]];
        double ar = tri_area(P[i], a, b);
        s = s + ar * (P[i] + a + b);
      }
      Pn[i] = unit(s);
    }
    P.swap(Pn);
    circumcentres();
  }

  // ---- checks: Delaunay and well-centred --------------------------------
  for (int i = 0; i < nC; i++) {
    int deg = ring_off[i + 1] - ring_off[i];
    for (int m = 0; m < deg; m++) {
      int t = ring_tri[ring_off[i] + m];
      int j = ring_nb[ring_off[i] + m], k = ring_nb[ring_off[i] + (m + 1) % deg];
      double rad = dot(P[i], cc[t]);
      int mprev = (m + deg - 1) % deg;
      int l = ring_nb[ring_off[i] + mprev];  // opposite point across edge (i,j)
      if (dot(P[l], cc[t]) > rad * (1 + 1e-12))
        throw std::runtime_error("triangulation is not Delaunay after Lloyd");
      if (tri_area(P[i], P[j], P[k]) <= 0)
        throw std::runtime_error("triangle orientation flipped");
      if (tri_area(P[i], P[j], cc[t]) <= 0 || tri_area(P[i], cc[t], P[k]) <= 0)
        throw std::runtime_error("triangle is not well centred");
    }
  }

  // ---- 3. numbering -------------------------------------------------------
  std::vector<int> c_new(nC), c_old(nC);
  std::iota(c_old.begin(), c_old.end(), 0);
  if (order == MG_ORDER_HILBERT) {
    std::vector<uint64_t> key(nC);
    for (int i = 0; i < nC; i++) key[i] = hilbert_key(P[i]);
    std::stable_sort(c_old.begin(), c_old.end(),
                     [&](int a, int b) { return key[a] < key[b]; });
  } else if (order != MG_ORDER_NATIVE) {
    throw std::runtime_error("unknown order");
  }
  for (int n = 0; n < nC; n++) c_new[c_old[n]] = n;

  // edges: (i, j) pairs, generation order = first appearance in rings
  struct EdgeRec {
    int a, b;      // new cell ids, a < b
    int ta, tb;    // old triangle ids on either side
    int64_t gen;   // generation order
  };
  std::vector<EdgeRec> edges;
  edges.reserve(size_t(nC) * 3);
  for (int i = 0; i < nC; i++) {
    int deg = ring_off[i + 1] - ring_off[i];
    for (int m = 0; m < deg; m++) {
      int j = ring_nb[ring_off[i] + m];
      if (j < i) continue;
      int t1 = ring_tri[ring_off[i] + m];
      int t0 = ring_tri[ring_off[i] + (m + deg - 1) % deg];
      int a = c_new[i], b = c_new[j];
      edges.push_back({std::min(a, b), std::max(a, b), t0, t1, int64_t(edges.size())});
    }
  }
  const int nE = int(edges.size());
  if (nE != 3 * (nC - 2)) throw std::runtime_error("unexpected edge count");
  if (order == MG_ORDER_HILBERT)
    std::sort(edges.begin(), edges.end(), [](const EdgeRec &x, const EdgeRec &y) {
      return x.a != y.a ? x.a < y.a : x.b < y.b;
    });

  std::vector<int> v_new(nV), v_old(nV);
  std::iota(v_old.begin(), v_old.end(), 0);
  std::vector<std::array<int, 3>> vcells(nV);
  for (int t = 0; t < nV; t++) {
    std::array<int, 3> s = {c_new[T[t].v[0]], c_new[T[t].v[1]], c_new[T[t].v[2]]};
    std::sort(s.begin(), s.end());
    vcells[t] = s;
  }
  if (order == MG_ORDER_HILBERT)
    std::stable_sort(v_old.begin(), v_old.end(),
                     [&](int x, int y) { return vcells[x] < vcells[y]; });
  for (int n = 0; n < nV; n++) v_new[v_old[n]] = n;

  // lookup (new cell a, new cell b) -> new edge id
  std::vector<std::vector<std::pair<int, int>>> edge_of(nC);
  for (int e = 0; e < nE; e++) edge_of[edges[e].a].push_back({edges[e].b, e});
  auto edge_id = [&](int a, int b) {
    int lo = std::min(a, b), hi = std::max(a, b);
    for (auto &pr : edge_of[lo])
      if (pr.first == hi) return pr.second;
    throw std::runtime_error("edge lookup failed");
  };

  // ---- 4. fields in the new numbering -----------------------------------
  M.nCells = nC;
  M.nEdges = nE;
  M.nVertices = nV;
  auto &I = M.ints;
  auto &D = M.dbls;
  auto ivec = [&](const char *n, size_t sz) -> std::vector<int32_t> & {
    auto &v = I[n];
    v.assign(sz, 0);
    return v;
  };
  auto dvec = [&](const char *n, size_t sz) -> std::vector<double> & {
    auto &v = D[n];
    v.assign(sz, 0.0);
    return v;
  };

  std::vector<V3> PC(nC), PV(nV), PE(nE);
  for (int n = 0; n < nC; n++) PC[n] = P[c_old[n]];
  for (int n = 0; n < nV; n++) PV[n] = cc[v_old[n]];

  auto set_latlon = [&](const std::vector<V3> &X, const char *px, const char *py,
                        const char *pz, const char *plat, const char *plon) {
    size_t n = X.size();
    auto &x = dvec(px, n);
    auto &y = dvec(py, n);
    auto &z = dvec(pz, n);
    auto &la = dvec(plat, n);
    auto &lo = dvec(plon, n);
    for (size_t i = 0; i < n; i++) {
      x[i] = radius * X[i].x;
      y[i] = radius * X[i].y;
      z[i] = radius * X[i].z;
      la[i] = std::asin(std::max(-1.0, std::min(1.0, X[i].z)));
      double l = std::atan2(X[i].y, X[i].x);
      if (l < 0) l += 2 * pi;
      lo[i] = l;
    }
  };

  // edges
  auto &cellsOnEdge = ivec("cellsOnEdge", size_t(nE) * 2);
  auto &verticesOnEdge = ivec("verticesOnEdge", size_t(nE) * 2);
  auto &dcEdge = dvec("dcEdge", nE);
  auto &dvEdge = dvec("dvEdge", nE);
  auto &angleEdge = dvec("angleEdge", nE);
  std::vector<int> ev1(nE), ev2(nE);
  for (int e = 0; e < nE; e++) {
    const EdgeRec &r = edges[e];
    V3 p1 = PC[r.a], p2 = PC[r.b];
    V3 xe = unit(p1 + p2);
    PE[e] = xe;
    V3 n = unit(p2 - p1);
    V3 t = cross(xe, n);
    int va = v_new[r.ta], vb = v_new[r.tb];
    if (dot(PV[vb] - PV[va], t) < 0) std::swap(va, vb);
    ev1[e] = va;
    ev2[e] = vb;
    cellsOnEdge[2 * e] = r.a + 1;
    cellsOnEdge[2 * e + 1] = r.b + 1;
    verticesOnEdge[2 * e] = va + 1;
    verticesOnEdge[2 * e + 1] = vb + 1;
    dcEdge[e] = radius * arc(p1, p2);
    dvEdge[e] = radius * arc(PV[va], PV[vb]);
    double rxy = std::sqrt(xe.x * xe.x + xe.y * xe.y);
    V3 east = {-xe.y / rxy, xe.x / rxy, 0.0};
    V3 north = cross(xe, east);
    angleEdge[e] = std::atan2(dot(n, north), dot(n, east));
  }

  // cells
  auto &nEdgesOnCell = ivec("nEdgesOnCell", nC);
  auto &edgesOnCell = ivec("edgesOnCell", size_t(nC) * kMaxEdges);
  auto &cellsOnCell = ivec("cellsOnCell", size_t(nC) * kMaxEdges);
  auto &verticesOnCell = ivec("verticesOnCell", size_t(nC) * kMaxEdges);
  for (int n = 0; n < nC; n++) {
    int i = c_old[n];
    int deg = ring_off[i + 1] - ring_off[i];
    // start the counter-clockwise walk at the lowest-numbered neighbour
    int m0 = 0;
    for (int m = 1; m < deg; m++)
      if (c_new[ring_nb[ring_off[i] + m]] < c_new[ring_nb[ring_off[i] + m0]]) m0 = m;
    nEdgesOnCell[n] = deg;
    for (int q = 0; q < deg; q++) {
      int m = (m0 + q) % deg;
      int j = c_new[ring_nb[ring_off[i] + m]];
      cellsOnCell[size_t(n) * kMaxEdges + q] = j + 1;
      edgesOnCell[size_t(n) * kMaxEdges + q] = edge_id(n, j) + 1;
      verticesOnCell[size_t(n) * kMaxEdges + q] = v_new[ring_tri[ring_off[i] + m]] + 1;
    }
  }

  // vertices
  auto &cellsOnVertex = ivec("cellsOnVertex", size_t(nV) * 3);
  auto &edgesOnVertex = ivec("edgesOnVertex", size_t(nV) * 3);
  auto &kite = dvec("kiteAreasOnVertex", size_t(nV) * 3);
  auto &areaTriangle = dvec("areaTriangle", nV);
  const double a2 = radius * radius;
  for (int v = 0; v < nV; v++) {
    auto cs = vcells[v_old[v]];
    V3 pa = PC[cs[0]], pb = PC[cs[1]], pc = PC[cs[2]];
    double ar = tri_area(pa, pb, pc);
    if (ar < 0) ar = tri_area(pa, pc, pb);
    areaTriangle[v] = a2 * ar;
    std::array<int, 3> es = {edge_id(cs[0], cs[1]), edge_id(cs[1], cs[2]),
                             edge_id(cs[0], cs[2])};
    std::sort(es.begin(), es.end());
    for (int k = 0; k < 3; k++) {
      cellsOnVertex[3 * v + k] = cs[k] + 1;
      edgesOnVertex[3 * v + k] = es[k] + 1;
    }
  }
  // kites, walking each cell's ring so the kite is traversed
  // counter-clockwise: (cell, x_e(j), vertex(j), x_e(j+1))
  auto &areaCell = dvec("areaCell", nC);
  for (int c = 0; c < nC; c++) {
    int deg = nEdgesOnCell[c];
    double acc = 0.0;
    for (int q = 0; q < deg; q++) {
      int e0 = edgesOnCell[size_t(c) * kMaxEdges + q] - 1;
      int e1 = edgesOnCell[size_t(c) * kMaxEdges + (q + 1) % deg] - 1;
      int v = verticesOnCell[size_t(c) * kMaxEdges + q] - 1;
      double k = tri_area(PC[c], PE[e0], PV[v]) + tri_area(PC[c], PV[v], PE[e1]);
      if (!(k > 0)) throw std::runtime_error("non-positive kite area");
      int pos = -1;
      for (int s = 0; s < 3; s++)
        if (cellsOnVertex[3 * v + s] - 1 == c) pos = s;
      if (pos < 0) throw std::runtime_error("cell/vertex mismatch");
      kite[3 * v + pos] = a2 * k;
      acc += a2 * k;
    }
    areaCell[c] = acc;
  }

  // edgesOnEdge / weightsOnEdge
  auto &nEdgesOnEdge = ivec("nEdgesOnEdge", nE);
  auto &edgesOnEdge = ivec("edgesOnEdge", size_t(nE) * kMaxEdges2);
  auto &weightsOnEdge = dvec("weightsOnEdge", size_t(nE) * kMaxEdges2);
  for (int e = 0; e < nE; e++) {
    int cnt = 0;
    for (int side = 0; side < 2; side++) {
      int c = cellsOnEdge[2 * e + side] - 1;
      int deg = nEdgesOnCell[c];
      int j0 = -1;
      for (int q = 0; q < deg; q++)
        if (edgesOnCell[size_t(c) * kMaxEdges + q] - 1 == e) j0 = q;
      if (j0 < 0) throw std::runtime_error("edge/cell mismatch");
      double n_e = (cellsOnEdge[2 * e] - 1 == c) ? 1.0 : -1.0;
      double R = 0.0;
      for (int m = 1; m < deg; m++) {
        int vpass = verticesOnCell[size_t(c) * kMaxEdges + (j0 + m - 1) % deg] - 1;
        int pos = -1;
        for (int s = 0; s < 3; s++)
          if (cellsOnVertex[3 * vpass + s] - 1 == c) pos = s;
        R += kite[3 * vpass + pos] / areaCell[c];
        int ep = edgesOnCell[size_t(c) * kMaxEdges + (j0 + m) % deg] - 1;
        double n_ep = (cellsOnEdge[2 * ep] - 1 == c) ? 1.0 : -1.0;
        edgesOnEdge[size_t(e) * kMaxEdges2 + cnt] = ep + 1;
        weightsOnEdge[size_t(e) * kMaxEdges2 + cnt] =
            n_e * n_ep * (0.5 - R) * dvEdge[ep] / dcEdge[e];
        cnt++;
      }
    }
    nEdgesOnEdge[e] = cnt;
  }

  set_latlon(PC, "xCell", "yCell", "zCell", "latCell", "lonCell");
  set_latlon(PE, "xEdge", "yEdge", "zEdge", "latEdge", "lonEdge");
  set_latlon(PV, "xVertex", "yVertex", "zVertex", "latVertex", "lonVertex");
  auto &meshDensity = dvec("meshDensity", nC);
  std::fill(meshDensity.begin(), meshDensity.end(), 1.0);
  auto &fV = dvec("fVertex", nV);
  auto &fE = dvec("fEdge", nE);
  auto &fC = dvec("fCell", nC);
  for (int v = 0; v < nV; v++) fV[v] = 2.0 * omega * std::sin(D["latVertex"][v]);
  for (int e = 0; e < nE; e++) fE[e] = 2.0 * omega * std::sin(D["latEdge"][e]);
  for (int c = 0; c < nC; c++) fC[c] = 2.0 * omega * std::sin(D["latCell"][c]);
}

}  // namespace

extern "C" {

mg_mesh *mg_generate(int level, double radius, int lloyd_iters, int order,
                     double omega) {
  return mg_generate_oriented(level, radius, lloyd_iters, order, omega, 1);
}

mg_mesh *mg_generate_oriented(int level, double radius, int lloyd_iters, int order,
                              double omega, int rotate) {
  try {
    auto *m = new mg_mesh();
    try {
      build(*m, level, radius, lloyd_iters, order, omega, rotate);
    } catch (...) {
      delete m;
      throw;
    }
    return m;
  } catch (const std::exception &ex) {
    g_err = ex.what();
    return nullptr;
  }
}

void mg_free(mg_mesh *m) { delete m; }

const char *mg_last_error(void) { return g_err.c_str(); }

int64_t mg_dim(const mg_mesh *m, const char *name) {
  if (!m || !name) return -1;
  std::string n(name);
  if (n == "nCells") return m->nCells;
  if (n == "nEdges") return m->nEdges;
  if (n == "nVertices") return m->nVertices;
  if (n == "maxEdges") return kMaxEdges;
  if (n == "maxEdges2") return kMaxEdges2;
  if (n == "vertexDegree") return 3;
  return -1;
}

const int32_t *mg_int_field(const mg_mesh *m, const char *name, int64_t *n) {
  if (!m || !name) return nullptr;
  auto it = m->ints.find(name);
  if (it == m->ints.end()) return nullptr;
  if (n) *n = int64_t(it->second.size());
  return it->second.data();
}

const double *mg_dbl_field(const mg_mesh *m, const char *name, int64_t *n) {
  if (!m || !name) return nullptr;
  auto it = m->dbls.find(name);
  if (it == m->dbls.end()) return nullptr;
  if (n) *n = int64_t(it->second.size());
  return it->second.data();
}

}  // extern "C"
