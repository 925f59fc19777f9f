// Context lifetime, device layouts of the mesh, and the device-resident
// shallow-water state (the counterpart of mesh::initialize plus the field
// storage FleCSI owns in the reference, mpasoflecsi/specialization/mesh.cc:67-116,
// apps/sw/state.hh:74-87).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "ctx.h"
#include "stage.h"

namespace mpasx {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string &msg) { g_err = msg; }

int check_ctx(const mpas_ctx *ctx) {
  if (!ctx) MPAS_FAIL(MPAS_EINVAL, "context is NULL");
  MPAS_HIP(hipSetDevice(ctx->device));
  return 0;
}

int after_launch(const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) MPAS_FAIL(MPAS_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

int check_finite(const double *p, size_t n, const char *what) {
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(p[i]))
      MPAS_FAIL(MPAS_ENONFINITE, std::string(what) + "[" + std::to_string(i) + "] is " +
                                     (std::isnan(p[i]) ? "NaN" : "infinite"));
  return 0;
}

// Counts the non-finite values of p[0, n): one atomic per wave that holds any.
__global__ __launch_bounds__(256) void k_count_nonfinite(const double *__restrict__ p, size_t n,
                                                         unsigned long long *count) {
  for (size_t i0 = size_t(blockIdx.x) * blockDim.x; i0 < n; i0 += size_t(gridDim.x) * blockDim.x) {
    const size_t i = i0 + threadIdx.x;
    const bool bad = i < n && !__builtin_isfinite(p[i]);
    const unsigned long long b = __ballot(bad);
    if ((threadIdx.x & 63u) == 0 && b) atomicAdd(count, (unsigned long long)__popcll(b));
  }
}

namespace {

// A slot-major ELL array [j][E] of width pad (even) as slot pairs,
// [j/2][E][2].
template <class T>
std::vector<T> slot_pairs(const std::vector<T> &a, int pad, int E) {
  std::vector<T> out(a.size());
  for (int p = 0; p < pad / 2; p++)
    for (int e = 0; e < E; e++)
      for (int s = 0; s < 2; s++)
        out[(size_t(p) * E + e) * 2 + s] = a[size_t(2 * p + s) * E + e];
  return out;
}

template <class T>
int upload(mpas_ctx *ctx, const std::vector<T> &v, const T **out) {
  *out = nullptr;
  if (v.empty()) return 0;
  void *p = nullptr;
  if (hipMalloc(&p, sizeof(T) * v.size()) != hipSuccess)
    MPAS_FAIL(MPAS_ENOMEM, "hipMalloc of " + std::to_string(sizeof(T) * v.size()) +
                               " bytes for the mesh failed");
  ctx->mesh_bufs.push_back(p);
  MPAS_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  *out = static_cast<const T *>(p);
  return 0;
}

// Cell staging lists of the one-level cell kernels, for workgroups of cpb
// cells: per block its own cells in order, then the other cells across
// their edges, ascending (list, off); loc[c] holds in bits 10 j .. the
// position of the cell across edge slot j of cell c (padding slots: its
// own).  False when a list would not fit 10-bit positions or `bytes_per`
// bytes of LDS per listed cell within 64 KiB.
struct CellLists {
  std::vector<int32_t> off, list;
  std::vector<uint64_t> loc;
  int mx = 0;
};
bool cell_lists(const HostMesh &m, int cpb, size_t bytes_per, CellLists &out) {
  const int C = m.nC;
  const int nb = (C + cpb - 1) / cpb;
  out.off.assign(nb + 1, 0);
  out.list.clear();
  out.loc.assign(C, 0);
  out.mx = 0;
  std::vector<int32_t> halo;
  auto other = [&](int c, int j) {
    const int e = m.c2e.at(c, j);
    return m.e2c.at(e, 0) == c ? m.e2c.at(e, 1) : m.e2c.at(e, 0);
  };
  for (int b = 0; b < nb; b++) {
    const int c0 = b * cpb, c1e = std::min(C, c0 + cpb);
    halo.clear();
    for (int c = c0; c < c1e; c++) {
      if (m.c2e.size(c) > 6) return false;  // six 10-bit positions per cell
      for (int j = 0; j < m.c2e.size(c); j++) {
        const int o = other(c, j);
        if (o < c0 || o >= c1e) halo.push_back(o);
      }
    }
    std::sort(halo.begin(), halo.end());
    halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
    const int own = c1e - c0, len = own + int(halo.size());
    if (len >= 1024 || size_t(len) * bytes_per > (64u << 10)) return false;
    out.mx = std::max(out.mx, len);
    for (int c = c0; c < c1e; c++) {
      uint64_t w = 0;
      for (int j = 0; j < 6; j++) {
        uint64_t p = uint64_t(c - c0);
        if (j < m.c2e.size(c)) {
          const int o = other(c, j);
          p = (o >= c0 && o < c1e)
                  ? uint64_t(o - c0)
                  : uint64_t(own + (std::lower_bound(halo.begin(), halo.end(), o) - halo.begin()));
        }
        w |= p << (10 * j);
      }
      out.loc[c] = w;
    }
    for (int c = c0; c < c1e; c++) out.list.push_back(c);
    out.list.insert(out.list.end(), halo.begin(), halo.end());
    out.off[b + 1] = int32_t(out.list.size());
  }
  return true;
}

int alloc_field(mpas_ctx *ctx, size_t n, double **out) {
  *out = nullptr;
  if (!n) return 0;
  void *p = nullptr;
  if (hipMalloc(&p, sizeof(double) * n) != hipSuccess)
    MPAS_FAIL(MPAS_ENOMEM, "hipMalloc of " + std::to_string(sizeof(double) * n) +
                               " bytes for the state failed");
  ctx->mesh_bufs.push_back(p);
  MPAS_HIP(zero_now(p, sizeof(double) * n));
  *out = static_cast<double *>(p);
  return 0;
}

// Builds every device array of the DMesh view from the flattened host mesh.
int build_device_mesh(mpas_ctx *ctx) {
  const HostMesh &m = ctx->host;
  DMesh &M = ctx->dm;
  const int C = m.nC, E = m.nE, V = m.nV;
  M.nC = C;
  M.nE = E;
  M.nV = V;
  M.nL = m.nL;
  M.nT = ctx->params.nTracers;
  M.gravity = ctx->params.gravity;
  M.coef3rd = ctx->params.coef3rd;
  M.wC = max_row(m.c2e);
  M.wEoE = max_row(m.e2e);
  if (M.wC > 8) MPAS_FAIL(MPAS_EINVAL, "cells with more than 8 edges are not supported");
  if (M.wEoE > 16) MPAS_FAIL(MPAS_EINVAL, "edgesOnEdge rows hold at most 16 entries");
  M.wAdv = 0;
  if (m.fct)
    for (int e = 0; e < E; e++) M.wAdv = std::max(M.wAdv, int(m.nAdvCellsForEdge[e]));
  // MPAS_HIP_WIDE_KERNELS=1 selects the 8 / 16 widths on any mesh (they are
  // otherwise used only by meshes with heptagons or wider stencils)
  const char *wide = std::getenv("MPAS_HIP_WIDE_KERNELS");
  M.narrow = M.wC <= 6 && M.wEoE <= 10 && M.wAdv <= 10 && !(wide && wide[0] == '1');
  // ELL widths as allocated: the kernels' compile-time widths
  const int padC = M.narrow ? 6 : 8, padE = M.narrow ? 10 : 16;

  int rc = 0;
  std::vector<int2> e2c(E), e2v(E);
  for (int e = 0; e < E; e++) {
    e2c[e] = make_int2(m.e2c.at(e, 0), m.e2c.at(e, 1));
    e2v[e] = make_int2(m.e2v.at(e, 0), m.e2v.at(e, 1));
  }
  if ((rc = upload(ctx, e2c, &M.e2c)) || (rc = upload(ctx, e2v, &M.e2v)) ||
      (rc = upload(ctx, m.dcEdge, &M.dc)) || (rc = upload(ctx, m.dvEdge, &M.dv)))
    return rc;
  {
    std::vector<uint8_t> n(E);
    std::vector<double> w(size_t(padE) * E, 0.0);
    for (int e = 0; e < E; e++) {
      n[e] = uint8_t(m.e2e.size(e));
      for (int j = 0; j < m.e2e.size(e); j++) w[size_t(j) * E + e] = m.weightsOnEdge[size_t(e) * 16 + j];
    }
    const auto ids = to_ell(m.e2e, padE);
    if ((rc = upload(ctx, n, &M.nEoE)) || (rc = upload(ctx, ids, &M.eoe)) ||
        (rc = upload(ctx, w, &M.woe)) || (rc = upload(ctx, slot_pairs(ids, padE, E), &M.eoeP)) ||
        (rc = upload(ctx, slot_pairs(w, padE, E), &M.woeP)))
      return rc;
    // LDS staging of the one-level momentum kernel (MPAS_HIP_EDGES_LDS=0
    // turns it off): per workgroup the distinct edgesOnEdge ids of its
    // edges, each edge's slots as positions in that list.  The kernel's
    // gathers of u, pv_edge and h_edge at ten slots touch about fifty cache
    // lines per wave-wide load; the staging copy reads each list in
    // ascending, mostly consecutive ids.
    const char *stg = std::getenv("MPAS_HIP_EDGES_LDS");
    if (m.nL == 1 && padE <= 12 && !(stg && stg[0] == '0')) {
      const int nb = (E + kEdgesBlock - 1) / kEdgesBlock;
      std::vector<int32_t> off(nb + 1, 0), list;
      std::vector<uint64_t> loc(size_t(2) * E, 0);
      int mx = 0;
      bool ok = true;
      std::vector<int32_t> es;
      for (int b = 0; b < nb && ok; b++) {
        const int e0 = b * kEdgesBlock, e1 = std::min(E, e0 + kEdgesBlock);
        es.clear();
        for (int e = e0; e < e1; e++)
          for (int j = 0; j < m.e2e.size(e); j++) es.push_back(m.e2e.at(e, j));
        std::sort(es.begin(), es.end());
        es.erase(std::unique(es.begin(), es.end()), es.end());
        // positions take 10 bits; the copy (3 doubles per entry) stays
        // within 24 KiB
        if (es.size() >= 1024) {
          ok = false;
          break;
        }
        mx = std::max(mx, int(es.size()));
        for (int e = e0; e < e1; e++)
          for (int j = 0; j < m.e2e.size(e); j++) {
            const uint64_t pos =
                uint64_t(std::lower_bound(es.begin(), es.end(), m.e2e.at(e, j)) - es.begin());
            loc[size_t(2) * e + j / 6] |= pos << (10 * (j % 6));
          }
        list.insert(list.end(), es.begin(), es.end());
        off[b + 1] = int32_t(list.size());
      }
      if (ok) {
        if ((rc = upload(ctx, loc, &M.e2eLoc)) || (rc = upload(ctx, off, &M.eStgOff)) ||
            (rc = upload(ctx, list, &M.eStgEdge)))
          return rc;
        M.eStgMax = mx;
        M.eStgLen = int(list.size());
      }
    }
  }
  // cells
  {
    std::vector<uint8_t> n(C), neg(C, 0);
    std::vector<uint32_t> perm(C, 0), rank(C, 0xffffffffu);
    Csr opp;
    opp.off = m.c2e.off;
    opp.idx.resize(m.c2e.idx.size());
    bool aligned = true;
    for (int c = 0; c < C; c++) {
      const int ne = m.c2e.size(c);
      n[c] = uint8_t(ne);
      int slot[8];
      for (int j = 0; j < ne; j++) {
        const int e = m.c2e.at(c, j);
        const int c1 = m.e2c.at(e, 0), c2 = m.e2c.at(e, 1);
        if (c == c1) neg[c] |= uint8_t(1u << j);
        opp.idx[m.c2e.off[c] + j] = (c == c1) ? c2 : c1;
        if (j >= m.c2c.size(c) || m.c2c.at(c, j) != opp.idx[m.c2e.off[c] + j]) aligned = false;
        slot[j] = j;
      }
      // ascending edge id: the original ids of a reordered mesh
      auto key = [&](int j) {
        const int e = m.c2e.at(c, j);
        return m.edgeOf.empty() ? e : m.edgeOf[e];
      };
      std::sort(slot, slot + ne, [&](int a, int b) { return key(a) < key(b); });
      for (int r = 0; r < ne; r++) {
        perm[c] |= uint32_t(slot[r]) << (4 * r);
        rank[c] &= ~(0xfu << (4 * slot[r]));
        rank[c] |= uint32_t(r) << (4 * slot[r]);
      }
      if (m.c2c.size(c) != ne) aligned = false;
    }
    ctx->aligned_c2c = aligned;
    if ((rc = upload(ctx, m.areaCell, &M.area)) || (rc = upload(ctx, n, &M.nEoC)) ||
        (rc = upload(ctx, neg, &M.sgnNeg)) || (rc = upload(ctx, perm, &M.ascPerm)) ||
        (rc = upload(ctx, rank, &M.ascRank)) ||
        (rc = upload(ctx, to_ell(m.c2e, padC), &M.c2e)))
      return rc;
    if (aligned) {
      if ((rc = upload(ctx, to_ell(m.c2c, padC), &M.c2c))) return rc;
      M.cOpp = M.c2c;  // cellsOnCell[j] lies across edgesOnCell[j]: one array serves both
    } else {
      if ((rc = upload(ctx, to_ell(opp, padC), &M.cOpp))) return rc;
      int wcc = std::max(1, max_row(m.c2c));
      if (wcc > 8) MPAS_FAIL(MPAS_EINVAL, "cellsOnCell rows hold at most 8 entries");
      if ((rc = upload(ctx, to_ell(m.c2c, std::max(wcc, padC)), &M.c2c))) return rc;
    }
  }
  // vertices
  {
    std::vector<double> kite(size_t(3) * V);
    std::vector<int32_t> v2e(size_t(3) * V);
    for (int v = 0; v < V; v++)
      for (int i = 0; i < 3; i++) {
        kite[size_t(i) * V + v] = m.kite[size_t(v) * 3 + i];
        if (m.v2e.size(v) != 3)
          MPAS_FAIL(MPAS_EINVAL, "vertex " + std::to_string(v) + " has " +
                                     std::to_string(m.v2e.size(v)) + " edges (expected 3)");
        const int e = m.v2e.at(v, i);
        const bool first = m.e2v.at(e, 0) == v;
        v2e[size_t(i) * V + v] = int32_t(uint32_t(e) | (first ? 0x80000000u : 0u));
      }
    if ((rc = upload(ctx, m.areaTriangle, &M.areaTri)) || (rc = upload(ctx, m.fVertex, &M.fV)) ||
        (rc = upload(ctx, kite, &M.kite)) || (rc = upload(ctx, to_ell(m.v2c, 3), &M.v2c)) ||
        (rc = upload(ctx, v2e, &M.v2e)))
      return rc;
  }
  // FCT / del2 setup products
  M.msd = m.fct ? m.meshScalingDel2[0] : 1.0;
  if (m.fct) {
    std::vector<uint8_t> na(E);
    std::vector<int32_t> adv(size_t(padE) * E);
    std::vector<double> c1(size_t(padE) * E, 0.0), c3(size_t(padE) * E, 0.0);
    for (int e = 0; e < E; e++) {
      na[e] = uint8_t(m.nAdvCellsForEdge[e]);
      for (int i = 0; i < padE; i++) {
        const bool live = i < m.nAdvCellsForEdge[e];
        adv[size_t(i) * E + e] = live ? m.advCellsForEdge[size_t(e) * 16 + i] : m.e2c.at(e, 0);
        c1[size_t(i) * E + e] = live ? m.advCoef[size_t(e) * 16 + i] : 0.0;
        c3[size_t(i) * E + e] = live ? m.advCoef_3rd[size_t(e) * 16 + i] : 0.0;
      }
    }
    // init_tracer_adv_coeff lists the edge's two cells first
    M.advC12 = 1;
    for (int e = 0; e < E && M.advC12; e++)
      if (na[e] < 2 || adv[e] != m.e2c.at(e, 0) || adv[size_t(E) + e] != m.e2c.at(e, 1))
        M.advC12 = 0;
    if ((rc = upload(ctx, na, &M.nAdv)) || (rc = upload(ctx, adv, &M.adv)) ||
        (rc = upload(ctx, c1, &M.advCoef)) || (rc = upload(ctx, c3, &M.advCoef3)) ||
        (rc = upload(ctx, slot_pairs(adv, padE, E), &M.advP)) ||
        (rc = upload(ctx, slot_pairs(c1, padE, E), &M.advCoefP)) ||
        (rc = upload(ctx, slot_pairs(c3, padE, E), &M.advCoef3P)))
      return rc;
    // advCoef_3rd in compact form: a slot that only one of the edge's cell
    // fits holds has advCoef_3rd = +-advCoef exactly (init_functions.cc:
    // 394-437 forms both from the same single derivTwo term); the code is
    // taken from the bits themselves, so the kernel reproduces every value
    {
      std::vector<uint32_t> side(E, 0);
      std::vector<double> x(size_t(kAdvExplicit) * E, 0.0);
      bool fits = M.narrow;
      for (int e = 0; e < E && fits; e++) {
        int k = 0;
        for (int i = 0; i < padE; i++) {
          const double a = c1[size_t(i) * E + e], b = c3[size_t(i) * E + e];
          uint32_t code;
          if (std::memcmp(&a, &b, sizeof a) == 0) {
            code = 1;
          } else {
            const double na_ = -a;
            if (std::memcmp(&na_, &b, sizeof a) == 0) {
              code = 2;
            } else {
              code = 0;
              if (k == kAdvExplicit) {
                fits = false;
                break;
              }
              x[size_t(k++) * E + e] = b;
            }
          }
          side[e] |= code << (2 * i);
        }
      }
      if (fits && ((rc = upload(ctx, side, &M.advSide)) ||
                   (rc = upload(ctx, slot_pairs(x, kAdvExplicit, E), &M.advC3xP))))
        return rc;
    }
    // LDS staging of the one-level high-order flux kernel (MPAS_HIP_HO_LDS=0
    // turns it off): per workgroup the distinct advection cells of its
    // edges, each edge's slots as positions in that list (as for the
    // momentum kernel; the tracers of a slot are one 16- or 32-byte record)
    {
      const char *hs = std::getenv("MPAS_HIP_HO_LDS");
      const int nT = ctx->params.nTracers;
      const bool full_cols = std::all_of(m.maxLevelCell.begin(), m.maxLevelCell.end(),
                                         [&](int x) { return x == m.nL; });
      const int npl = nT > 0 ? ho_lanes_per_edge(nT) : 1;
      if (m.nL == 1 && nT > 0 && M.narrow && padE <= 12 && M.advSide && full_cols &&
          kHoBlock % npl == 0 && !(hs && hs[0] == '0')) {
        const int epb = kHoBlock / npl;
        const int nb = (E + epb - 1) / epb;
        std::vector<int32_t> off(nb + 1, 0), list, cs;
        std::vector<uint64_t> loc(size_t(2) * E, 0);
        int mx = 0;
        bool ok = true;
        for (int b = 0; b < nb && ok; b++) {
          const int e0 = b * epb, e1 = std::min(E, e0 + epb);
          cs.clear();
          for (int e = e0; e < e1; e++)
            for (int i = 0; i < padE; i++) cs.push_back(adv[size_t(i) * E + e]);
          std::sort(cs.begin(), cs.end());
          cs.erase(std::unique(cs.begin(), cs.end()), cs.end());
          // 10-bit positions; the copy stays within 64 KiB
          if (cs.size() >= 1024 || cs.size() * nT * sizeof(double) > (64u << 10)) {
            ok = false;
            break;
          }
          mx = std::max(mx, int(cs.size()));
          for (int e = e0; e < e1; e++)
            for (int i = 0; i < padE; i++) {
              const uint64_t pos = uint64_t(
                  std::lower_bound(cs.begin(), cs.end(), adv[size_t(i) * E + e]) - cs.begin());
              loc[size_t(2) * e + i / 6] |= pos << (10 * (i % 6));
            }
          list.insert(list.end(), cs.begin(), cs.end());
          off[b + 1] = int32_t(list.size());
        }
        if (ok) {
          if ((rc = upload(ctx, loc, &M.a2cLoc)) || (rc = upload(ctx, off, &M.aStgOff)) ||
              (rc = upload(ctx, list, &M.aStgCell)))
            return rc;
          M.aStgMax = mx;
          M.aStgLen = int(list.size());
          M.aStgEpb = epb;
        }
      }
    }
    // LDS staging of the one-level cells kernel (FCT), on request
    // (MPAS_HIP_CELLS_LDS=1): per workgroup the distinct edges of its cells,
    // each cell's slots as positions in them.  Off by default: on the
    // 655,362-cell mesh with 4 tracers the kernel takes 81 us with it and
    // 71 us without, since L2 already serves the repeated edge reads and
    // the copy adds a barrier-separated pass per workgroup.
    const char *stg = std::getenv("MPAS_HIP_CELLS_LDS");
    const int nT = ctx->params.nTracers;
    if (m.nL == 1 && nT > 0 && M.narrow && ctx->params.tracer_scheme == MPAS_TRACER_FCT &&
        stg && stg[0] == '1') {
      const int npl = nT % 2 == 0 ? nT / 2 : nT;  // tracer lanes per cell
      const int cpb = kCellsBlock / npl;
      if (kCellsBlock % npl == 0) {
        const int nb = (C + cpb - 1) / cpb;
        std::vector<int32_t> off(nb + 1, 0), list;
        std::vector<uint64_t> loc(C, 0);
        int mx = 0;
        bool ok = true;
        for (int b = 0; b < nb && ok; b++) {
          const int c0 = b * cpb, c1e = std::min(C, c0 + cpb);
          std::vector<int32_t> es;
          for (int c = c0; c < c1e; c++) {
            ok = ok && m.c2e.size(c) <= 6;  // six 10-bit positions per cell
            for (int j = 0; j < m.c2e.size(c); j++) es.push_back(m.c2e.at(c, j));
          }
          std::sort(es.begin(), es.end());
          es.erase(std::unique(es.begin(), es.end()), es.end());
          // positions take 10 bits; the LDS copy stays within 64 KiB
          if (es.size() >= 1024 || es.size() * (5 + nT) * sizeof(double) > (64u << 10)) ok = false;
          mx = std::max(mx, int(es.size()));
          for (int c = c0; c < c1e; c++)
            for (int j = 0; j < m.c2e.size(c); j++) {
              const uint64_t pos = uint64_t(std::lower_bound(es.begin(), es.end(), m.c2e.at(c, j)) -
                                            es.begin());
              loc[c] |= pos << (10 * j);
            }
          list.insert(list.end(), es.begin(), es.end());
          off[b + 1] = int32_t(list.size());
        }
        if (ok && ((rc = upload(ctx, loc, &M.c2eLoc)) || (rc = upload(ctx, off, &M.stgOff)) ||
                   (rc = upload(ctx, list, &M.stgEdge))))
          return rc;
        if (ok) {
          M.stgMax = mx;
          M.stgCpb = cpb;
        }
      }
    }
    // LDS staging of cell values in the one-level tracer update
    // (MPAS_HIP_TRACER_LDS=0 turns it off: its cells' and their
    // neighbours' limiter factors and tracers) and, on request,
    // the cells kernel (MPAS_HIP_CELLS_CST=1: their tracers and thickness),
    // each over its own workgroups' cells.  The cells kernel is slower
    // with it on the 655,362-cell mesh with 4 tracers (84.0 against 72.9 us
    // per launch): at 118 VGPRs it holds 4 waves per SIMD, and the copy's
    // barrier costs more than the gathers it saves.
    const bool full_cols = std::all_of(m.maxLevelCell.begin(), m.maxLevelCell.end(),
                                       [&](int x) { return x == m.nL; });
    if (m.nL == 1 && nT > 0 && nT % 2 == 0 && M.narrow && full_cols &&
        ctx->params.tracer_scheme == MPAS_TRACER_FCT) {
      const char *ts = std::getenv("MPAS_HIP_TRACER_LDS");
      const int tnpl = tracer_lanes_per_cell(nT);
      CellLists L;
      if (kTracerBlock % tnpl == 0 && !(ts && ts[0] == '0') &&
          cell_lists(m, kTracerBlock / tnpl, 3 * nT * sizeof(double), L)) {
        if ((rc = upload(ctx, L.loc, &M.c2cLoc)) || (rc = upload(ctx, L.off, &M.tStgOff)) ||
            (rc = upload(ctx, L.list, &M.tStgCell)))
          return rc;
        M.tStgMax = L.mx;
        M.tStgLen = int(L.list.size());
        M.tStgCpb = kTracerBlock / tnpl;
      }
      const char *cs = std::getenv("MPAS_HIP_CELLS_CST");
      const int cnpl = nT / 2;  // the cells kernel takes tracer pairs
      if (kCellsBlock % cnpl == 0 && cs && cs[0] == '1' && !M.c2eLoc &&
          cell_lists(m, kCellsBlock / cnpl, (nT + 1) * sizeof(double), L)) {
        if ((rc = upload(ctx, L.loc, &M.c2cLocC)) || (rc = upload(ctx, L.off, &M.cStgOff)) ||
            (rc = upload(ctx, L.list, &M.cStgCell)))
          return rc;
        M.cStgMax = L.mx;
        M.cStgLen = int(L.list.size());
        M.cStgCpb = kCellsBlock / cnpl;
      }
    }
    bool full = std::all_of(m.maxLevelCell.begin(), m.maxLevelCell.end(),
                            [&](int x) { return x == m.nL; });
    if (!full) {
      if ((rc = upload(ctx, m.maxLevelCell, &M.mlc)) ||
          (rc = upload(ctx, m.maxLevelEdgeTop, &M.mlet)))
        return rc;
    }
    bool all_ho = std::all_of(m.highOrderAdvMask.begin(), m.highOrderAdvMask.end(),
                              [](int x) { return x == 1; });
    if (!all_ho && (rc = upload(ctx, m.highOrderAdvMask, &M.hoMask))) return rc;
  }
  return 0;
}

int release(mpas_ctx *ctx) {
  adj_free(ctx);
  for (hipEvent_t e : ctx->prof.pool) (void)hipEventDestroy(e);
  ctx->prof.pool.clear();
  for (void *p : ctx->mesh_bufs) (void)hipFree(p);
  ctx->mesh_bufs.clear();
  for (void *p : ctx->user_bufs) (void)hipFree(p);
  ctx->user_bufs.clear();
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
  ctx->own_stream = nullptr;
  return 0;
}

int alloc_state(mpas_ctx *ctx) {
  if (ctx->sw) return 0;
  auto W = std::make_unique<SwState>();
  const DMesh &M = ctx->dm;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, V = size_t(M.nV) * M.nL;
  const size_t T = C * M.nT;
  int rc = 0;
  double **fields_e[] = {&W->u, &W->nu, &W->pu[0], &W->pu[1], &W->h_edge[0], &W->h_edge[1],
                         &W->pv_edge[0], &W->pv_edge[1]};
  double **fields_c[] = {&W->h, &W->nh, &W->ph[0], &W->ph[1], &W->ke[0], &W->ke[1],
                         &W->div[0], &W->div[1]};
  double **fields_t[] = {&W->tr, &W->ntr, &W->ptr[0], &W->ptr[1], &W->tlo};
  for (auto f : fields_e)
    if ((rc = alloc_field(ctx, E, f))) return rc;
  for (auto f : fields_c)
    if ((rc = alloc_field(ctx, C, f))) return rc;
  for (auto f : fields_t)
    if ((rc = alloc_field(ctx, T, f))) return rc;
  if ((rc = alloc_field(ctx, V, &W->vort[0])) || (rc = alloc_field(ctx, V, &W->vort[1])) ||
      (rc = alloc_field(ctx, V, &W->pv_vertex)) || (rc = alloc_field(ctx, size_t(M.nC), &W->h_s)) ||
      (rc = alloc_field(ctx, E * M.nT, &W->anti)) || (rc = alloc_field(ctx, 2 * T, &W->rfac)))
    return rc;
  double *cnt = nullptr;
  if ((rc = alloc_field(ctx, 1, &cnt))) return rc;
  W->nonfinite = reinterpret_cast<unsigned long long *>(cnt);
  ctx->sw = std::move(W);
  return 0;
}

}  // namespace
}  // namespace mpasx

using namespace mpasx;

extern "C" {

const char *mpas_last_error(void) { return g_err.c_str(); }

static int prepare_host_mesh(const mpas_mesh_desc *mesh, const mpas_params *params,
                             HostMesh &hm) {
  if (!mesh || !params) MPAS_FAIL(MPAS_EINVAL, "mesh or params is NULL");
  const mpas_params &p = *params;
  if (p.nLevels < 1 || p.nLevels > 64) MPAS_FAIL(MPAS_EINVAL, "nLevels must be in [1,64]");
  if (p.nTracers < 0 || p.nTracers > 8) MPAS_FAIL(MPAS_EINVAL, "nTracers must be in [0,8]");
  if (p.tracer_scheme < MPAS_TRACER_NONE || p.tracer_scheme > MPAS_TRACER_FCT)
    MPAS_FAIL(MPAS_EINVAL, "unknown tracer_scheme");
  if (p.tracer_scheme != MPAS_TRACER_NONE && p.nTracers == 0)
    MPAS_FAIL(MPAS_EINVAL, "a tracer scheme needs nTracers > 0");
  if (!(p.gravity > 0)) MPAS_FAIL(MPAS_EINVAL, "gravity must be positive");
  std::string err = flatten_mesh(*mesh, p.nLevels, hm);
  if (!err.empty()) MPAS_FAIL(MPAS_EINVAL, "mesh rejected: " + err);
  const bool need_setup = p.tracer_scheme == MPAS_TRACER_FCT || p.kappa2 != 0.0;
  if (need_setup) {
    if (!(p.sphere_radius > 0)) MPAS_FAIL(MPAS_EINVAL, "sphere_radius must be positive");
    err = setup_tracer_operators(hm, p.sphere_radius);
    if (!err.empty()) MPAS_FAIL(MPAS_EINVAL, "FCT setup failed: " + err);
  }
  return 0;
}

int mpas_host_mesh_create(const mpas_mesh_desc *mesh, const mpas_params *params,
                          mpas_host_mesh **out) {
  if (!out) MPAS_FAIL(MPAS_EINVAL, "mpas_host_mesh_create: out is NULL");
  *out = nullptr;
  auto hm = std::make_unique<mpas_host_mesh>();
  if (int rc = prepare_host_mesh(mesh, params, *hm)) return rc;
  *out = hm.release();
  return 0;
}

int mpas_host_mesh_reorder(const mpas_host_mesh *hm, int order, mpas_host_mesh **out) {
  if (!hm || !out) MPAS_FAIL(MPAS_EINVAL, "mpas_host_mesh_reorder: NULL argument");
  *out = nullptr;
  auto o = std::make_unique<mpas_host_mesh>();
  const std::string err = reorder_host_mesh(*hm, order, *o);
  if (!err.empty()) MPAS_FAIL(MPAS_EINVAL, "mpas_host_mesh_reorder: " + err);
  *out = o.release();
  return 0;
}

int mpas_host_mesh_order(const mpas_host_mesh *hm, int32_t *cell_of, int32_t *edge_of,
                         int32_t *vertex_of) {
  if (!hm) MPAS_FAIL(MPAS_EINVAL, "mpas_host_mesh_order: NULL mesh");
  auto put = [](const std::vector<int32_t> &of, int n, int32_t *dst) {
    if (!dst) return;
    for (int i = 0; i < n; i++) dst[i] = of.empty() ? i : of[i];
  };
  put(hm->cellOf, hm->nC, cell_of);
  put(hm->edgeOf, hm->nE, edge_of);
  put(hm->vertexOf, hm->nV, vertex_of);
  return 0;
}

int mpas_host_mesh_destroy(mpas_host_mesh *hm) {
  delete hm;
  return 0;
}

const mpas_host_mesh *mpas_ctx_host_mesh(const mpas_ctx *ctx) {
  return ctx ? &ctx->host : nullptr;
}

static int finish_ctx(std::unique_ptr<mpas_ctx> ctx, const mpas_params *params,
                      mpas_ctx **out);

int mpas_ctx_create(const mpas_mesh_desc *mesh, const mpas_params *params, mpas_ctx **out) {
  if (!out) MPAS_FAIL(MPAS_EINVAL, "mpas_ctx_create: out is NULL");
  *out = nullptr;
  auto ctx = std::make_unique<mpas_ctx>();
  if (int rc = prepare_host_mesh(mesh, params, ctx->host)) return rc;
  return finish_ctx(std::move(ctx), params, out);
}

int mpas_ctx_create_reordered(const mpas_ctx *src, int order, mpas_ctx **out) {
  if (!out) MPAS_FAIL(MPAS_EINVAL, "mpas_ctx_create_reordered: out is NULL");
  *out = nullptr;
  if (int rc = check_ctx(src)) return rc;
  if (order < MPAS_ORDER_GIVEN || order > MPAS_ORDER_MORTON)
    MPAS_FAIL(MPAS_EINVAL, "mpas_ctx_create_reordered: unknown order");
  auto ctx = std::make_unique<mpas_ctx>();
  const std::string err = reorder_host_mesh(src->host, order, ctx->host);
  if (!err.empty()) MPAS_FAIL(MPAS_EINVAL, "mpas_ctx_create_reordered: " + err);
  return finish_ctx(std::move(ctx), &src->params, out);
}

int mpas_ctx_order(const mpas_ctx *ctx, int32_t *cell_of, int32_t *edge_of, int32_t *vertex_of) {
  if (!ctx) MPAS_FAIL(MPAS_EINVAL, "context is NULL");
  return mpas_host_mesh_order(&ctx->host, cell_of, edge_of, vertex_of);
}

static int finish_ctx(std::unique_ptr<mpas_ctx> ctx, const mpas_params *params,
                      mpas_ctx **out) {
  ctx->params = *params;
  {
    // The kernels form byte offsets in 32 bits (ld / st): every array they
    // index must stay below 4 GiB.  The largest are the 16-wide edge lists
    // of doubles, the limiter factors [C][L][T][2] and the [E][L][T] fluxes.
    const HostMesh &h = ctx->host;
    const double nT = std::max(1, params->nTracers), nL = h.nL;
    const double worst = std::max({16.0 * 8 * h.nE, 16.0 * h.nC * nL * nT, 8.0 * h.nE * nL * nT,
                                   8.0 * h.nV * nL * 3});
    if (worst >= 4294967296.0)
      MPAS_FAIL(MPAS_ESHAPE, "mesh too large for one context: an array would exceed 4 GiB");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    MPAS_FAIL(MPAS_EHIP, "no HIP device is visible");
  MPAS_HIP(hipGetDevice(&ctx->device));
  MPAS_HIP(hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking));
  ctx->stream = ctx->own_stream;
  {
    const HostMesh &h = ctx->host;
    ctx->sum_eoe = double(h.e2e.idx.size());
    ctx->sum_eoc = double(h.c2e.idx.size());
    for (int x : h.nAdvCellsForEdge) ctx->sum_adv += x;
  }
  int rc = build_device_mesh(ctx.get());
  if (rc) {
    std::string keep = g_err;
    release(ctx.get());
    g_err = keep;
    return rc;
  }
  *out = ctx.release();
  return 0;
}

int mpas_ctx_destroy(mpas_ctx *ctx) {
  if (!ctx) return 0;
  (void)hipSetDevice(ctx->device);
  if (ctx->own_stream) (void)hipStreamSynchronize(ctx->own_stream);
  release(ctx);
  delete ctx;
  return 0;
}

int mpas_ctx_set_stream(mpas_ctx *ctx, void *hip_stream) {
  if (!ctx) MPAS_FAIL(MPAS_EINVAL, "context is NULL");
  ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own_stream;
  return 0;
}

int mpas_ctx_sync(mpas_ctx *ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int mpas_ctx_dims(const mpas_ctx *ctx, int64_t out[5]) {
  if (!ctx || !out) MPAS_FAIL(MPAS_EINVAL, "mpas_ctx_dims: NULL argument");
  out[0] = ctx->dm.nC;
  out[1] = ctx->dm.nE;
  out[2] = ctx->dm.nV;
  out[3] = ctx->dm.nL;
  out[4] = ctx->dm.nT;
  return 0;
}

int mpas_buf_alloc(mpas_ctx *ctx, size_t bytes, void **dptr) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!dptr) MPAS_FAIL(MPAS_EINVAL, "mpas_buf_alloc: dptr is NULL");
  *dptr = nullptr;
  if (!bytes) MPAS_FAIL(MPAS_EINVAL, "mpas_buf_alloc: zero bytes");
  if (hipMalloc(dptr, bytes) != hipSuccess)
    MPAS_FAIL(MPAS_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  ctx->user_bufs.insert(*dptr);
  return 0;
}

int mpas_buf_free(mpas_ctx *ctx, void *dptr) {
  if (int rc = check_ctx(ctx)) return rc;
  auto it = ctx->user_bufs.find(dptr);
  if (it == ctx->user_bufs.end()) MPAS_FAIL(MPAS_EINVAL, "mpas_buf_free: unknown buffer");
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  MPAS_HIP(hipFree(dptr));
  ctx->user_bufs.erase(it);
  return 0;
}

int mpas_buf_upload(mpas_ctx *ctx, void *dptr, const void *host, size_t bytes) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!dptr || !host) MPAS_FAIL(MPAS_EINVAL, "mpas_buf_upload: NULL pointer");
  MPAS_HIP(hipMemcpyAsync(dptr, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int mpas_buf_download(mpas_ctx *ctx, void *host, const void *dptr, size_t bytes) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!dptr || !host) MPAS_FAIL(MPAS_EINVAL, "mpas_buf_download: NULL pointer");
  MPAS_HIP(hipMemcpyAsync(host, dptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int mpas_mesh_connectivity(const mpas_host_mesh *hm, const char *which, int32_t *offsets,
                           int32_t *indices, int64_t *rows, int64_t *count) {
  if (!hm || !which || !rows || !count) MPAS_FAIL(MPAS_EINVAL, "mpas_mesh_connectivity: NULL");
  const HostMesh &m = *hm;
  const std::string w(which);
  const Csr *c = w == "c2v" ? &m.c2v : w == "c2e" ? &m.c2e : w == "c2c" ? &m.c2c
               : w == "e2v" ? &m.e2v : w == "e2e" ? &m.e2e : w == "v2c" ? &m.v2c
               : w == "e2c" ? &m.e2c : w == "v2e" ? &m.v2e : nullptr;
  if (!c) MPAS_FAIL(MPAS_EINVAL, "unknown connectivity '" + w + "'");
  *rows = c->rows();
  *count = int64_t(c->idx.size());
  if (offsets) std::memcpy(offsets, c->off.data(), sizeof(int32_t) * c->off.size());
  if (indices) std::memcpy(indices, c->idx.data(), sizeof(int32_t) * c->idx.size());
  return 0;
}

int mpas_mesh_setup_field(const mpas_host_mesh *hm, const char *name, void *host, int64_t n) {
  if (!hm || !name || !host) MPAS_FAIL(MPAS_EINVAL, "mpas_mesh_setup_field: NULL");
  const HostMesh &m = *hm;
  if (!m.fct) MPAS_FAIL(MPAS_ESTATE, "the context was created without the FCT setup");
  const std::string s(name);
  const std::vector<int32_t> *vi = s == "maxLevelEdgeTop" ? &m.maxLevelEdgeTop
                                 : s == "nAdvCellsForEdge" ? &m.nAdvCellsForEdge
                                 : s == "advCellsForEdge" ? &m.advCellsForEdge
                                 : s == "highOrderAdvMask" ? &m.highOrderAdvMask : nullptr;
  const std::vector<double> *vd = s == "edgeSignOnCell" ? &m.edgeSignOnCell
                                : s == "meshScalingDel2" ? &m.meshScalingDel2
                                : s == "derivTwo" ? &m.derivTwo
                                : s == "advCoef" ? &m.advCoef
                                : s == "advCoef_3rd" ? &m.advCoef_3rd : nullptr;
  if (vi) {
    if (int64_t(vi->size()) != n)
      MPAS_FAIL(MPAS_ESHAPE, s + " has " + std::to_string(vi->size()) + " entries");
    std::memcpy(host, vi->data(), sizeof(int32_t) * n);
    return 0;
  }
  if (vd) {
    if (int64_t(vd->size()) != n)
      MPAS_FAIL(MPAS_ESHAPE, s + " has " + std::to_string(vd->size()) + " entries");
    std::memcpy(host, vd->data(), sizeof(double) * n);
    return 0;
  }
  MPAS_FAIL(MPAS_EINVAL, "unknown setup field '" + s + "'");
}

int mpas_sw_state_upload(mpas_ctx *ctx, const double *h, const double *u, const double *tracers,
                         const double *h_s) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!h || !u || !h_s) MPAS_FAIL(MPAS_EINVAL, "mpas_sw_state_upload: h, u and h_s are required");
  const DMesh &M = ctx->dm;
  if (M.nT && !tracers) MPAS_FAIL(MPAS_EINVAL, "mpas_sw_state_upload: tracers is NULL");
  {
    const size_t C = size_t(M.nC) * M.nL, E = size_t(M.nE) * M.nL;
    int rc;
    if ((rc = check_finite(h, C, "mpas_sw_state_upload: h")) ||
        (rc = check_finite(u, E, "mpas_sw_state_upload: u")) ||
        (rc = check_finite(h_s, size_t(M.nC), "mpas_sw_state_upload: h_s")) ||
        (M.nT && (rc = check_finite(tracers, C * M.nT, "mpas_sw_state_upload: tracers"))))
      return rc;
  }
  if (int rc = alloc_state(ctx)) return rc;
  SwState &W = *ctx->sw;
  const size_t cb = sizeof(double) * M.nC * M.nL, eb = sizeof(double) * M.nE * M.nL;
  MPAS_HIP(hipMemcpyAsync(W.h, h, cb, hipMemcpyHostToDevice, ctx->stream));
  MPAS_HIP(hipMemcpyAsync(W.u, u, eb, hipMemcpyHostToDevice, ctx->stream));
  MPAS_HIP(hipMemcpyAsync(W.h_s, h_s, sizeof(double) * M.nC, hipMemcpyHostToDevice, ctx->stream));
  if (M.nT)
    MPAS_HIP(hipMemcpyAsync(W.tr, tracers, cb * M.nT, hipMemcpyHostToDevice, ctx->stream));
  W.cur = 0;
  W.elapsed = 0;
  // apps/sw/run.cc:88-92: diagnostics of the initial state
  if (int rc = sw_diagnostics(ctx, W.u, W.h, W.cur)) return rc;
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  W.ready = true;
  return 0;
}

int mpas_sw_state_check(mpas_ctx *ctx, int64_t *nonfinite) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->sw || !ctx->sw->ready) MPAS_FAIL(MPAS_ESTATE, "no state uploaded");
  const DMesh &M = ctx->dm;
  SwState &W = *ctx->sw;
  const size_t C = size_t(M.nC) * M.nL, E = size_t(M.nE) * M.nL;
  MPAS_HIP(hipMemsetAsync(W.nonfinite, 0, sizeof(unsigned long long), ctx->stream));
  const double *arr[3] = {W.h, W.u, W.tr};
  const size_t len[3] = {C, E, M.nT ? C * M.nT : 0};
  for (int a = 0; a < 3; a++) {
    if (!len[a]) continue;
    hipLaunchKernelGGL(k_count_nonfinite, dim3(std::min<size_t>((len[a] + 255) / 256, 4096)),
                       dim3(256), 0, ctx->stream, arr[a], len[a], W.nonfinite);
    if (int rc = after_launch("k_count_nonfinite")) return rc;
  }
  unsigned long long n = 0;
  MPAS_HIP(hipMemcpyAsync(&n, W.nonfinite, sizeof n, hipMemcpyDeviceToHost, ctx->stream));
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  if (nonfinite) *nonfinite = int64_t(n);
  if (n) MPAS_FAIL(MPAS_ENONFINITE, std::to_string(n) + " non-finite values in h, u and tracers");
  return 0;
}

int mpas_sw_state_download(mpas_ctx *ctx, double *h, double *u, double *tracers) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->sw || !ctx->sw->ready) MPAS_FAIL(MPAS_ESTATE, "no state uploaded");
  const DMesh &M = ctx->dm;
  SwState &W = *ctx->sw;
  const size_t cb = sizeof(double) * M.nC * M.nL, eb = sizeof(double) * M.nE * M.nL;
  if (h) MPAS_HIP(hipMemcpyAsync(h, W.h, cb, hipMemcpyDeviceToHost, ctx->stream));
  if (u) MPAS_HIP(hipMemcpyAsync(u, W.u, eb, hipMemcpyDeviceToHost, ctx->stream));
  if (tracers && M.nT)
    MPAS_HIP(hipMemcpyAsync(tracers, W.tr, cb * M.nT, hipMemcpyDeviceToHost, ctx->stream));
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int mpas_sw_field_download(mpas_ctx *ctx, const char *name, double *host, int64_t n) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!name || !host) MPAS_FAIL(MPAS_EINVAL, "mpas_sw_field_download: NULL");
  if (!ctx->sw || !ctx->sw->ready) MPAS_FAIL(MPAS_ESTATE, "no state uploaded");
  const DMesh &M = ctx->dm;
  SwState &W = *ctx->sw;
  const std::string s(name);
  const int64_t C = int64_t(M.nC) * M.nL, E = int64_t(M.nE) * M.nL, V = int64_t(M.nV) * M.nL;
  const double *src = nullptr;
  int64_t len = 0;
  if (s == "h") src = W.h, len = C;
  else if (s == "u") src = W.u, len = E;
  else if (s == "tracers") src = W.tr, len = C * M.nT;
  else if (s == "h_edge") src = W.h_edge[W.cur], len = E;
  else if (s == "ke") src = W.ke[W.cur], len = C;
  else if (s == "pv_edge") src = W.pv_edge[W.cur], len = E;
  else if (s == "vorticity") src = W.vort[W.cur], len = V;
  else if (s == "pv_vertex") src = W.pv_vertex, len = V;
  else if (s == "h_s") src = W.h_s, len = M.nC;
  else if (s == "divergence") src = W.div[W.cur], len = C;
  else MPAS_FAIL(MPAS_EINVAL, "unknown field '" + s + "'");
  if (n != len) MPAS_FAIL(MPAS_ESHAPE, s + " has " + std::to_string(len) + " entries");
  if (len) MPAS_HIP(hipMemcpyAsync(host, src, sizeof(double) * len, hipMemcpyDeviceToHost, ctx->stream));
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int mpas_rk4_step(mpas_ctx *ctx, double dt, int nsteps) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->sw || !ctx->sw->ready)
    MPAS_FAIL(MPAS_ESTATE, "mpas_rk4_step: upload a state first (mpas_sw_state_upload)");
  if (nsteps < 0) MPAS_FAIL(MPAS_EINVAL, "mpas_rk4_step: nsteps < 0");
  if (!(dt > 0)) MPAS_FAIL(MPAS_EINVAL, "mpas_rk4_step: dt must be positive");
  const DMesh &M = ctx->dm;
  if (M.nT && ctx->params.tracer_scheme == MPAS_TRACER_FCT) {
    if (M.mlc) MPAS_FAIL(MPAS_EINVAL,
                         "the fused FCT step needs every level active (maxLevelCell == nLevels); "
                         "use the per-task entry points for partial columns");
    if (!ctx->aligned_c2c)
      MPAS_FAIL(MPAS_EINVAL, "the fused FCT step needs cellsOnCell[j] across edgesOnCell[j]");
  }
  return sw_rk4_steps(ctx, dt, nsteps);
}

int mpas_ocean_rk4_step(mpas_ctx *ctx, double dt, int nsteps) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->sw || !ctx->sw->ready)
    MPAS_FAIL(MPAS_ESTATE, "mpas_ocean_rk4_step: upload a state first (mpas_sw_state_upload)");
  if (nsteps < 0) MPAS_FAIL(MPAS_EINVAL, "mpas_ocean_rk4_step: nsteps < 0");
  if (!(dt > 0)) MPAS_FAIL(MPAS_EINVAL, "mpas_ocean_rk4_step: dt must be positive");
  const DMesh &M = ctx->dm;
  if (!M.nT || ctx->params.tracer_scheme != MPAS_TRACER_FCT)
    MPAS_FAIL(MPAS_EINVAL, "the fused ocean run advects tracers with FCT (use the per-task "
                           "entry points for del2 mixing alone)");
  if (M.mlc || !ctx->aligned_c2c)
    MPAS_FAIL(MPAS_EINVAL, "the fused ocean run needs every level active and cellsOnCell[j] "
                           "across edgesOnCell[j]");
  return ocean_rk4_steps(ctx, dt, nsteps);
}

}  // extern "C"

// ---- per-kernel profiling ----------------------------------------------------
namespace mpasx {

static int prof_event(mpas_ctx *ctx, hipEvent_t *ev) {
  Prof &P = ctx->prof;
  if (P.used == P.pool.size()) {
    hipEvent_t e;
    MPAS_HIP(hipEventCreate(&e));
    P.pool.push_back(e);
  }
  *ev = P.pool[P.used++];
  return 0;
}

int prof_begin(mpas_ctx *ctx, const char *kernel, double bytes, hipStream_t st) {
  if (!ctx) return 0;
  Prof &P = ctx->prof;
  if (!P.on) return 0;
  if (!st) st = ctx->stream;
  Prof::Pending p{P.id(kernel), nullptr, nullptr, bytes, st};
  if (int rc = prof_event(ctx, &p.start)) return rc;
  if (int rc = prof_event(ctx, &p.stop)) return rc;
  MPAS_HIP(hipEventRecord(p.start, st));
  P.pending.push_back(p);
  return 0;
}

int prof_end(mpas_ctx *ctx, hipStream_t st) {
  if (!ctx) return 0;
  Prof &P = ctx->prof;
  if (!P.on) return 0;
  if (!st) st = ctx->stream;
  for (auto it = P.pending.rbegin(); it != P.pending.rend(); ++it)
    if (it->stream == st) {
      MPAS_HIP(hipEventRecord(it->stop, st));
      return 0;
    }
  MPAS_FAIL(MPAS_EINVAL, "prof_end without prof_begin");
}

}  // namespace mpasx

extern "C" {

void *mpas_ctx_stream(const mpas_ctx *ctx) { return ctx ? ctx->stream : nullptr; }

int mpas_prof_enable(mpas_ctx *ctx, int on) {
  if (int rc = check_ctx(ctx)) return rc;
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  ctx->prof.on = on != 0;
  ctx->prof.pending.clear();
  ctx->prof.used = 0;
  return 0;
}

int mpas_prof_report(mpas_ctx *ctx, char *buf, size_t cap) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!buf || !cap) MPAS_FAIL(MPAS_EINVAL, "mpas_prof_report: empty buffer");
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  Prof &P = ctx->prof;
  std::vector<double> ms(P.names.size(), 0.0), bytes(P.names.size(), 0.0);
  std::vector<long> n(P.names.size(), 0);
  for (auto &p : P.pending) {
    float t = 0;
    MPAS_HIP(hipEventElapsedTime(&t, p.start, p.stop));
    ms[p.kernel] += t;
    bytes[p.kernel] += p.bytes;
    n[p.kernel]++;
  }
  std::string out = "[";
  for (size_t i = 0; i < P.names.size(); i++) {
    if (!n[i]) continue;
    char line[256];
    std::snprintf(line, sizeof line,
                  "%s{\"kernel\": \"%s\", \"launches\": %ld, \"total_ms\": %.6f, "
                  "\"bytes_per_launch\": %.1f}",
                  out.size() > 1 ? ", " : "", P.names[i].c_str(), n[i], ms[i], bytes[i] / n[i]);
    out += line;
  }
  out += "]";
  if (out.size() + 1 > cap) MPAS_FAIL(MPAS_ESHAPE, "mpas_prof_report: buffer too small");
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return 0;
}

}  // extern "C"
