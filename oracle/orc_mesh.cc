// ORACLE — test infrastructure only (see orc_internal.h).
// Topology flattening restated from the reference.
#include <stdexcept>

#include "orc_internal.h"

namespace orc {

// mpasoflecsi/specialization/definition.hh:670-718: rows of a 2-D MPAS
// connectivity dataset become CRS lists; zero entries are dropped and the
// 1-based file ids become 0-based.
void read_connectivity_2D(const int32_t *buf, size_t dim1, size_t dim2, Crs &data) {
  data.off.clear();
  data.idx.clear();
  data.off.push_back(0);
  size_t counter = 0;
  for (size_t i = 0; i < dim1; i++) {
    for (size_t j = 0; j < dim2; j++) {
      if (buf[i * dim2 + j] != 0) {
        data.idx.push_back(buf[i * dim2 + j] - 1);
        counter++;
      }
    }
    data.off.push_back(int(counter));
  }
}

// mpasoflecsi/specialization/mesh.cc:93-105: copy the file order.
void init_cnx(const Crs &cnx, Crs &out) {
  out.off.assign(1, 0);
  out.idx.clear();
  for (size_t c = 0; c + 1 < cnx.off.size(); ++c) {
    auto start = cnx.off[c];
    auto size = cnx.off[c + 1] - start;
    for (int i = 0; i < size; ++i) out.idx.push_back(cnx.idx[start + i]);
    out.off.push_back(int(out.idx.size()));
  }
}

// mpasoflecsi/specialization/mesh.cc:108-116: v2c[c2v[c][v]].push_back(c),
// so every transposed list is in ascending source id.
void transpose_cnx(const Crs &c2v, int n_to, Crs &v2c) {
  std::vector<std::vector<int>> rag(n_to);
  for (int c = 0; c < c2v.rows(); ++c)
    for (int v = 0; v < c2v.size(c); ++v) rag.at(c2v.at(c, v)).push_back(c);
  v2c.off.assign(1, 0);
  v2c.idx.clear();
  for (int i = 0; i < n_to; i++) {
    for (int x : rag[i]) v2c.idx.push_back(x);
    v2c.off.push_back(int(v2c.idx.size()));
  }
}

// mpasoflecsi/specialization/definition.hh:1079-1092 (which datasets) and
// mpasoflecsi/specialization/mesh.cc:67-90 (which are copied, which are
// transposed).
void finalize_mesh(Mesh &m) {
  Crs f_c2v, f_c2e, f_c2c, f_e2v, f_e2e;
  read_connectivity_2D(m.f_voc.data(), m.nC, m.maxEdges, f_c2v);
  read_connectivity_2D(m.f_eoc.data(), m.nC, m.maxEdges, f_c2e);
  read_connectivity_2D(m.f_coc.data(), m.nC, m.maxEdges, f_c2c);
  read_connectivity_2D(m.f_voe.data(), m.nE, 2, f_e2v);
  read_connectivity_2D(m.f_eoe.data(), m.nE, m.maxEdges2, f_e2e);

  init_cnx(f_c2v, m.c2v);
  transpose_cnx(m.c2v, m.nV, m.v2c);
  init_cnx(f_c2e, m.c2e);
  transpose_cnx(m.c2e, m.nE, m.e2c);
  init_cnx(f_c2c, m.c2c);
  init_cnx(f_e2v, m.e2v);
  transpose_cnx(m.e2v, m.nV, m.v2e);
  init_cnx(f_e2e, m.e2e);
}

}  // namespace orc
