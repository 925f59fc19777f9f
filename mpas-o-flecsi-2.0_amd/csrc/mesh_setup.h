// Host-side, one-time mesh preparation of the MI355X library:
//  * topology flattening with the reference's orders
//    (mpasoflecsi/specialization/mesh.cc:67-116, definition.hh:670-718);
//  * the FCT / del2 setup products
//    (apps/ocean/tasks/init_functions.cc:18-447);
//  * the device layouts (int32 ELL, slot-major) built from them.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "mpas_hip.h"

namespace mpasx {

// Ragged list per entity (the FleCSI ragged connectivity field).
struct Csr {
  std::vector<int32_t> off{0};
  std::vector<int32_t> idx;
  int rows() const { return int(off.size()) - 1; }
  int size(int i) const { return off[i + 1] - off[i]; }
  int at(int i, int j) const { return idx[off[i] + j]; }
};

struct HostMesh {
  int nC = 0, nE = 0, nV = 0, nL = 1;
  int maxEdges = 8, maxEdges2 = 16, vdeg = 3;
  // connectivity exactly as mesh::initialize leaves it
  Csr c2v, c2e, c2c, e2v, e2e, v2c, e2c, v2e;
  std::vector<double> areaCell, dvEdge, dcEdge, areaTriangle, kite, fVertex,
      weightsOnEdge /*[E][maxEdges2]*/, xCell, yCell, zCell, xVertex, yVertex,
      zVertex, meshDensity;
  std::vector<int32_t> maxLevelCell /*[C]*/, boundaryCell /*[C][nL]*/;
  // setup products in the reference layouts (init_functions.cc)
  bool fct = false;
  std::vector<int32_t> maxLevelEdgeTop, nAdvCellsForEdge,
      advCellsForEdge /*[E][16]*/, highOrderAdvMask /*[E][nL]*/;
  std::vector<double> edgeSignOnCell /*[C][8]*/, meshScalingDel2 /*[E]*/,
      derivTwo /*[E][16][2]*/, advCoef /*[E][16]*/, advCoef_3rd /*[E][16]*/;
  // After a reordering: the original id of each entity (empty: the ids are
  // the original ones).  edgeKey orders a cell's edges for the sums the
  // reference forms in ascending edge id, so that a reordered mesh keeps
  // every sum's order and its bits.
  std::vector<int32_t> cellOf, edgeOf, vertexOf;
};

// Cell orders of reorder_host_mesh.
enum { kOrderGiven = 0, kOrderRcm = 1, kOrderMorton = 2, kOrders = 3 };

// The mesh m renumbered: cells in the given order (as m, reverse
// Cuthill-McKee, or along a 3-D Morton curve of the cell centres), edges by
// their (lower, higher) new cell pair, vertices by their sorted new cell
// triple.  Every list keeps its entries' order (ids renumbered), every
// per-entity array moves with its entity, and the original ids are kept
// (cellOf / edgeOf / vertexOf, composed with m's own).
std::string reorder_host_mesh(const HostMesh &m, int order, HostMesh &out);

// Returns an empty string on success, else the reason the mesh is rejected.
std::string flatten_mesh(const mpas_mesh_desc &d, int nLevels, HostMesh &m);
std::string setup_tracer_operators(HostMesh &m, double sphere_radius);

// Slot-major ELL copy of a ragged list: out[j * rows + i] = row i, entry j;
// padding repeats the row's own id (never read: kernels stop at the count).
std::vector<int32_t> to_ell(const Csr &c, int width);
int max_row(const Csr &c);

}  // namespace mpasx
