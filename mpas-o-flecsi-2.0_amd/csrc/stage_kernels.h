// Device bodies of the stage-fused RK4 kernels, one work item each: the
// single-instance kernels (fused_kernels.hip) and the ensemble kernels
// (batch.hip) wrap the same code, so every member of an ensemble is
// advanced with exactly the operations of a single run.
#pragma once
#include "ops.h"
#include "stage.h"

namespace mpasx {

template <int NT>
__device__ __forceinline__ void load_tr(const double *__restrict__ p, uint32_t i0,
                                        double (&v)[NT]) {
#pragma unroll
  for (int t = 0; t < NT; t++) v[t] = ld(p, i0 + t);
}

// Tracers handled by one lane: pairs when the count is even.
template <int NT>
constexpr int tracers_per_lane() {
  return NT % 2 == 0 ? 2 : 1;
}

// Adds the live slots' values in ascending edge id, the order of the
// reference's scatter loops: at step r the slot of rank r is selected with
// compile-time indices only (a run-time index into a register array would
// send the array to scratch memory).  acc -= v where the cell is the edge's
// first cell (negative edge sign), acc += v otherwise.
template <int W>
__device__ __forceinline__ double sum_ascending(double acc, const double (&v)[W],
                                                uint32_t rank, uint8_t neg, int n) {
#pragma unroll
  for (int r = 0; r < W; r++) {
    double f = 0.0;
    bool ng = false;
#pragma unroll
    for (int j = 0; j < W; j++) {
      const bool hit = ((rank >> (4 * j)) & 15) == uint32_t(r);
      f = hit ? v[j] : f;
      ng = hit ? bool((neg >> j) & 1) : ng;
    }
    const double x = ng ? acc - f : acc + f;
    acc = (r < n) ? x : acc;
  }
  return acc;
}

// A phase boundary for the instruction scheduler: the loads written before
// it issue together (index loads, then the gathers they feed) instead of
// being interleaved with their uses, which would serialise the memory round
// trips (the one-level kernels are one basic block, and the scheduler
// otherwise trades that latency for registers).
#ifndef MPAS_SCHED_PHASES
#define MPAS_SCHED_PHASES 1
#endif
#if MPAS_SCHED_PHASES
#define MPAS_PHASE() __builtin_amdgcn_sched_barrier(0)
#else
#define MPAS_PHASE() ((void)0)
#endif

// Static rows shared across an entity's lanes.  With levels on lanes the G
// lanes of an entity (G = chunks x tracer lane groups) would each load the
// same W slots of its stencil row: W vector loads per lane for one row.
// When G divides 64 the G lanes form an aligned block of a wave (an
// entity's lanes are adjacent items, entities start at multiples of G and
// workgroups are whole waves); if also G >= W, lane r of the block loads
// slot r only and every lane takes slot j from lane base + j by a
// cross-lane permute (ds_bpermute, no memory access).  Every lane of the
// block belongs to the same entity, so a lane that returned early (past
// the last entity) is never read.  The permutes run outside divergent
// code: an inactive source lane would not provide its value.
#ifndef MPAS_SHARE_ROWS
#define MPAS_SHARE_ROWS 1
#endif
template <int W>
__device__ __forceinline__ bool share_rows(int G) {
  return MPAS_SHARE_ROWS && G >= W && G <= 64 && (64 % G) == 0;
}
// This lane's slot (r, clamped to the row) and the block's first lane.
__device__ __forceinline__ int row_slot(int item, int G, int W, int &base) {
  const int r = item % G;
  base = int(threadIdx.x & 63u) - r;
  return r < W ? r : W - 1;
}
template <int W, class T>
__device__ __forceinline__ void bcast_row(T (&out)[W], T mine, int base) {
#pragma unroll
  for (int j = 0; j < W; j++) out[j] = __shfl(mine, base + j, 64);
}

// One-level kernels share a cell's rows across its G tracer lanes (an
// aligned group of G lanes, G dividing 64): out[j] = slot j, which the
// group's lane j % G holds as mine[j / G], moved by ds_bpermute.  (DPP quad
// permutations, for G = 2 and 4, measured no faster: cells 71.2 against
// 72.7 us per launch at level 8.)
template <int G, int W, int MS, class T>
__device__ __forceinline__ void grp_row(T (&out)[W], const T (&mine)[MS], int base) {
#pragma unroll
  for (int j = 0; j < W; j++) out[j] = __shfl(mine[j / G], base + j % G, 64);
}

// One-level cells kernel: the cell's tracer lanes share the new-u and
// dcEdge loads of the kinetic energy and divergence, which only the lead
// lane forms.
#ifndef MPAS_CELLS_SHARE_KE
#define MPAS_CELLS_SHARE_KE 1
#endif

// The one-level momentum and high-order flux kernels read their stencils
// from the slot-paired copies (one 8- or 16-byte access per slot pair).
#ifndef MPAS_PAIRED_STENCIL
#define MPAS_PAIRED_STENCIL 1
#endif

// The one-level high-order flux kernel reads advCoef_3rd in compact form
// (DMesh::advSide) where the context built it.
#ifndef MPAS_COMPACT_C3
#define MPAS_COMPACT_C3 1
#endif

// Slots of the momentum kernel's edgesOnEdge gathers per scheduling phase
// with levels on lanes (0: all at once).
#ifndef MPAS_QCH
#define MPAS_QCH 5
#endif

// The stage kernels read every slot of a ragged list up front, with widths
// fixed at compile time (WE edgesOnEdge, WA advection cells, W edges of a
// cell): all index loads issue together, then all gathers, so a lane waits
// for two memory round trips instead of two per slot.  Slots past a row's
// count read the ELL padding (valid ids) and are masked out of every sum,
// which still runs over the live slots in the reference's order.
//
// Tracer work runs NPL = NT / TPL lanes per entity, each lane handling TPL
// consecutive tracers (two when NT is even: one 16-byte access of the
// [cell][level][tracer] record per pair); the lanes of an entity are
// adjacent and the first of them writes the entity's non-tracer outputs.

// Levels of a work item.  LOL = 0 serves one-level contexts: a lane takes
// an entity's one level (k0 = 0, k1 = 1, so the compiler sees a single
// pass and keeps the 32-bit offset addressing; a level loop would be
// strength-reduced into 64-bit pointers held in registers).  Contexts of
// more levels put levels on lanes.  With levels on lanes (LOL = 1) a lane takes
// S.lpl consecutive levels of one entity, and the lanes of an entity cover
// its S.nchunk chunks of levels side by side: when the levels are ensemble
// members (every path kernel treats levels independently), neighbouring
// lanes then read neighbouring members of the same entity, so every state
// access and gather is one contiguous run.  Item order: entity, then level
// chunk, then tracer lane group.  LOL = 2 pads each entity's lanes to a
// multiple of the wave size, so all lanes of a wave share one entity: the
// entity index is then wave-uniform (taken from the wave's first lane) and
// every static mesh load becomes a scalar load; padding lanes get no
// levels.  FIX > 0: the launch guarantees S.lpl == FIX (and FIX divides the
// level count), so a lane's range is one pass of the level loop that the
// compiler can see (it then keeps the 32-bit offset addressing and fewer
// registers, as for one level); FIX = 0: S.lpl levels, the last chunk
// possibly shorter.
template <int LOL, int NPL, int FIX = 0>
__device__ __forceinline__ int split_levels(const StageIO &S, uint32_t L, int item, int &sub,
                                            uint32_t &k0, uint32_t &k1) {
  if constexpr (LOL == 0) {
    const int q = item / NPL;
    sub = item - q * NPL;
    k0 = 0;
    k1 = 1;
    return q;
  } else if constexpr (LOL == 1) {
    const int q = item / NPL;
    sub = item - q * NPL;
    const int ent = q / S.nchunk, ch = q - ent * S.nchunk;
    if constexpr (FIX > 0) {
      k0 = uint32_t(ch) * uint32_t(FIX);
      k1 = k0 + uint32_t(FIX);
    } else {
      k0 = uint32_t(ch) * uint32_t(S.lpl);
      k1 = min(k0 + uint32_t(S.lpl), L);
    }
    return ent;
  } else {
    const int pe = (S.nchunk * NPL + 63) / 64 * 64;
    const int ent = __builtin_amdgcn_readfirstlane(item) / pe;
    const int r = item - ent * pe, q = r / NPL;
    sub = r - q * NPL;
    if constexpr (FIX > 0) {
      k0 = uint32_t(q) * uint32_t(FIX);
      k1 = q < S.nchunk ? k0 + uint32_t(FIX) : k0;
    } else {
      k0 = uint32_t(q) * uint32_t(S.lpl);
      k1 = q < S.nchunk ? min(k0 + uint32_t(S.lpl), L) : k0;
    }
    return ent;
  }
}

// ---- edges: momentum ---------------------------------------------------------
// compute_tend (+ viscosity), compute_substep, accumulate_update and, at the
// last stage, finalize_timestep for u (solve.cc:215-233, 285-289, 326-330,
// 359-366)
//
// STG (one-level launches of kEdgesBlock lanes, DMesh::e2eLoc set): the
// workgroup first copies u, pv_edge and h_edge of the distinct edgesOnEdge
// edges of its edges into LDS (`lds`: 3 arrays of M.eStgMax doubles), then
// every edge reads its slots there by position.
template <int WE, int LOL = 0, int LP = 1, int FIX = 0, int QCX = MPAS_QCH, bool STG = false>
__device__ __forceinline__ void stage_edges(const DMesh &M, const StageIO &S, int item,
                                            double *lds = nullptr) {
  static_assert(!STG || (LOL == 0 && LP == 1 && WE <= 12), "LDS staging serves one-level launches");
  int sub;
  uint32_t k0, k1;
  const uint32_t mx = STG ? uint32_t(M.eStgMax) : 0u;
  double *const s_u = lds, *const s_pv = lds + mx, *const s_he = lds + 2 * mx;
  if constexpr (STG) {
    if (!S.freeze_u) {  // uniform: the whole grid skips the copy or takes part
      const int lb = item / kEdgesBlock;
      const int nb = (M.nE + kEdgesBlock - 1) / kEdgesBlock;
      if (lb < nb) {
        const int b0 = ld(M.eStgOff, uint32_t(lb)), ne = ld(M.eStgOff, uint32_t(lb + 1)) - b0;
        for (int i = int(threadIdx.x); i < ne; i += int(blockDim.x)) {
          const uint32_t x = uint32_t(ld(M.eStgEdge, uint32_t(b0 + i)));
          s_u[i] = ld(S.xu, x);
          s_pv[i] = ld(S.pv_edge, x);
          s_he[i] = ld(S.h_edge, x);
        }
      }
      __syncthreads();
    }
  }
  const int e = split_levels<LOL, 1, FIX>(S, M.nL, item, sub, k0, k1);
  if (e >= M.nE) return;
  const uint32_t L = M.nL, E = M.nE;
  if (S.freeze_u) {
    if (S.stage < 3)
      for (uint32_t k = k0; k < k1; k++) st(S.yu, e * L + k, ld(S.u0, e * L + k));
    return;
  }
  const int2 cc = ldc(M.e2c, e);
  const int n = ldc(M.nEoE, e);
  uint32_t id[WE];
  double w[WE];
  if constexpr (STG) {
    // slot positions in the workgroup's list, six per word
    const mpas_u2v lw = ld_vec<kNtStencil, mpas_u2v>(M.e2eLoc, uint32_t(e));
#pragma unroll
    for (int j = 0; j < WE; j++)
      id[j] = uint32_t((j < 6 ? lw.x : lw.y) >> (10 * (j % 6))) & 1023u;
#pragma unroll
    for (int p = 0; p < WE / 2; p++) {
      const mpas_d2v wj = ld_vec<kNtStencil, mpas_d2v>(M.woeP, p * E + e);
      w[2 * p] = wj.x;
      w[2 * p + 1] = wj.y;
    }
  } else if (LOL == 1 && share_rows<WE>(S.nchunk)) {
    int base;
    const int j = row_slot(item, S.nchunk, WE, base);
    bcast_row<WE, uint32_t>(id, uint32_t(ld(M.eoe, j * E + e)), base);
    bcast_row<WE, double>(w, ld(M.woe, j * E + e), base);
  } else if (LOL == 0 && MPAS_PAIRED_STENCIL && M.eoeP) {
    // one access per slot pair
#pragma unroll
    for (int p = 0; p < WE / 2; p++) {
      const mpas_i2v ij = ld_vec<kNtStencil, mpas_i2v>(M.eoeP, p * E + e);
      const mpas_d2v wj = ld_vec<kNtStencil, mpas_d2v>(M.woeP, p * E + e);
      id[2 * p] = uint32_t(ij.x);
      id[2 * p + 1] = uint32_t(ij.y);
      w[2 * p] = wj.x;
      w[2 * p + 1] = wj.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < WE; j++) {
      id[j] = ldn<LOL == 0 && kNtStencil>(M.eoe, j * E + e);
      w[j] = ldn<LOL == 0 && kNtStencil>(M.woe, j * E + e);
    }
  }
  const double dc = ldc(M.dc, e);
  for (uint32_t k = k0; k < k1; k += LP) {
    const uint32_t o = e * L + k;
    const uint32_t o1 = uint32_t(cc.x) * L + k, o2 = uint32_t(cc.y) * L + k;
    double pve[LP], ke1[LP], ke2[LP], xh1[LP], xh2[LP];
    ld_lv<LP>(S.pv_edge, o, pve);
    ld_lv<LP>(S.ke, o1, ke1);
    ld_lv<LP>(S.ke, o2, ke2);
    ld_lv<LP>(S.xh, o1, xh1);
    ld_lv<LP>(S.xh, o2, xh2);
    double visc[LP];
    if (S.use_visc) {
      const int2 vv = ldc(M.e2v, e);
      double d1[LP], d2[LP], v1[LP], v2[LP];
      ld_lv<LP>(S.div, o1, d1);
      ld_lv<LP>(S.div, o2, d2);
      ld_lv<LP>(S.vort, uint32_t(vv.x) * L + k, v1);
      ld_lv<LP>(S.vort, uint32_t(vv.y) * L + k, v2);
      const double dv = ldc(M.dv, e);
#pragma unroll
      for (int p = 0; p < LP; p++)
        visc[p] = S.visc_coef * ((d2[p] - d1[p]) / dc - (v2[p] - v1[p]) / dv);  // visc_at
    }
    constexpr bool NTS = LOL == 0 && kNtNext;
    double u0[LP], nu[LP];
    if constexpr (LP == 1) {
      if (S.stage < 3) u0[0] = ldo<NTS>(S.u0, o);
      if (S.stage > 0) nu[0] = ldo<NTS>(S.nu, o);
    } else {
      if (S.stage < 3) ld_lv<LP>(S.u0, o, u0);
      if (S.stage > 0) ld_lv<LP>(S.nu, o, nu);
    }
    // q over the edgesOnEdge slots in file order, gathered QCH slots at a
    // time (a scheduling phase each, so that the gathers of one chunk only
    // are live at once): with levels on lanes the freed registers buy
    // occupancy (level 7, 32 levels: 426 -> 409 us), one-level launches
    // were faster with every gather in flight at once (92 against 95 us)
    constexpr int QCH = LOL && QCX > 0 && QCX < WE ? QCX : WE;
    double qs[LP];
#pragma unroll
    for (int p = 0; p < LP; p++) qs[p] = 0.0;
#pragma unroll
    for (int j0 = 0; j0 < WE; j0 += QCH) {
      double uu[QCH][LP], pv[QCH][LP], he[QCH][LP];
#pragma unroll
      for (int jj = 0; jj < QCH && j0 + jj < WE; jj++) {
        if constexpr (!STG) {  // STG: read from LDS where the sum takes them
          const uint32_t oj = id[j0 + jj] * L + k;
          ld_lv<LP>(S.xu, oj, uu[jj]);
          ld_lv<LP>(S.pv_edge, oj, pv[jj]);
          ld_lv<LP>(S.h_edge, oj, he[jj]);
        }
      }
#pragma unroll
      for (int p = 0; p < LP; p++) {
#pragma unroll
        for (int jj = 0; jj < QCH && j0 + jj < WE; jj++) {
          const int j = j0 + jj;
          const double uj = STG ? s_u[id[j]] : uu[jj][p];
          const double pj = STG ? s_pv[id[j]] : pv[jj][p];
          const double hj = STG ? s_he[id[j]] : he[jj][p];
          const double workpv = 0.5 * (pve[p] + pj);
          const double qn = qs[p] + w[j] * uj * workpv * hj;
          qs[p] = (j < n) ? qn : qs[p];
        }
      }
      if (j0 + QCH < WE) MPAS_PHASE();
    }
    double yu[LP], nun[LP];
#pragma unroll
    for (int p = 0; p < LP; p++) {
      const double q = qs[p];
      double t = q - (ke2[p] - ke1[p] +
                      M.gravity * (xh2[p] + ldc(S.h_s, cc.y) - xh1[p] - ldc(S.h_s, cc.x))) /
                         dc;
      if (S.use_visc) t += visc[p];
      if (S.stage == 0) {
        yu[p] = u0[p] + S.w_sub * t;
        nun[p] = u0[p] + S.w_rk * t;
      } else if (S.stage < 3) {
        yu[p] = u0[p] + S.w_sub * t;
        nun[p] = nu[p] + S.w_rk * t;
      } else {
        yu[p] = nu[p] + S.w_rk * t;
      }
    }
    st_lv<LP>(S.yu, o, yu);
    if (S.stage < 3) {
      if constexpr (LP == 1) stn<NTS>(S.nu, o, nun[0]);
      else st_lv<LP>(S.nu, o, nun);
    }
  }
}

// ---- edges x tracers: FCT high-order minus low-order flux -------------------
// calc_normal_thickness_flux, the high-order flux and its correction by the
// low-order flux (ocean/solve.cc:40-57, 120-167, 216-229); TPL tracers per
// lane.
//
// STG (one-level launches of kHoBlock lanes, DMesh::a2cLoc set): the
// workgroup first copies the tracers of the distinct advection cells of
// its edges into LDS (`lds`: M.aStgMax records of NT doubles), then every
// edge reads its slots there by position.
template <int NT, int WA, int LOL = 0, int TPLX = 0, int LP = 1, int FIX = 0, bool STG = false,
          int QHX = 0>
__device__ __forceinline__ void stage_ho(const DMesh &M, const StageIO &S, int item,
                                         double *lds = nullptr) {
  constexpr int TPL = TPLX ? TPLX : tracers_per_lane<NT>(), NPL = NT / TPL;
  static_assert(!STG || (LOL == 0 && LP == 1 && WA <= 12), "LDS staging serves one-level launches");
  if constexpr (STG) {
    const int lb = item / kHoBlock;
    const int nb = (M.nE + M.aStgEpb - 1) / M.aStgEpb;
    if (lb < nb) {
      const int b0 = ld(M.aStgOff, uint32_t(lb)), nc = ld(M.aStgOff, uint32_t(lb + 1)) - b0;
      for (int i = int(threadIdx.x); i < nc; i += int(blockDim.x)) {
        const uint32_t c = uint32_t(ld(M.aStgCell, uint32_t(b0 + i)));
        double a[NT];
        ld_group<NT>(S.xtr, c * NT, a);
#pragma unroll
        for (int t = 0; t < NT; t++) lds[uint32_t(i) * NT + t] = a[t];
      }
    }
    __syncthreads();
  }
  int sub;
  uint32_t k0, k1;
  const int e = split_levels<LOL, NPL, FIX>(S, M.nL, item, sub, k0, k1);
  const int t0 = sub * TPL;
  if (e >= M.nE) return;
  const uint32_t L = M.nL, E = M.nE;
  const int2 cc = ldc(M.e2c, e);
  const double dv = ldc(M.dv, e);
  const int na = ldc(M.nAdv, e);
  const int lve = lvl_edge(M, e);
  uint32_t ic[WA];
  double c1[WA], c3[WA];
  int lvc[WA];
  if (LOL == 1 && share_rows<WA>(S.nchunk * NPL)) {
    int base;
    const int i = row_slot(item, S.nchunk * NPL, WA, base);
    bcast_row<WA, uint32_t>(ic, uint32_t(ld(M.adv, i * E + e)), base);
    bcast_row<WA, double>(c1, ld(M.advCoef, i * E + e), base);
    bcast_row<WA, double>(c3, ld(M.advCoef3, i * E + e), base);
#pragma unroll
    for (int j = 0; j < WA; j++) c3[j] = c3[j] * M.coef3rd;
  } else if (STG || (LOL == 0 && MPAS_PAIRED_STENCIL && MPAS_COMPACT_C3 && M.advP && M.advSide)) {
    // one access per slot pair; advCoef_3rd from its side code: +-advCoef,
    // or the edge's next explicit value (picked by selects: a run-time
    // index into a register array would go to scratch memory).  STG: the
    // slots' positions in the workgroup's cell list instead of cell ids.
    if constexpr (STG) {
      const mpas_u2v lw = ld_vec<kNtStencil, mpas_u2v>(M.a2cLoc, uint32_t(e));
#pragma unroll
      for (int i = 0; i < WA; i++)
        ic[i] = uint32_t((i < 6 ? lw.x : lw.y) >> (10 * (i % 6))) & 1023u;
    }
    const uint32_t side = ldn<kNtStencil>(M.advSide, e);
    double x[kAdvExplicit];
#pragma unroll
    for (int p = 0; p < kAdvExplicit / 2; p++) {
      const mpas_d2v v = ld_vec<kNtStencil, mpas_d2v>(M.advC3xP, p * E + e);
      x[2 * p] = v.x;
      x[2 * p + 1] = v.y;
    }
#pragma unroll
    for (int p = 0; p < WA / 2; p++) {
      if constexpr (!STG) {
        const mpas_i2v ij = ld_vec<kNtStencil, mpas_i2v>(M.advP, p * E + e);
        ic[2 * p] = uint32_t(ij.x);
        ic[2 * p + 1] = uint32_t(ij.y);
      }
      const mpas_d2v a1 = ld_vec<kNtStencil, mpas_d2v>(M.advCoefP, p * E + e);
      c1[2 * p] = a1.x;
      c1[2 * p + 1] = a1.y;
    }
    int next = 0;  // explicit values used so far
#pragma unroll
    for (int i = 0; i < WA; i++) {
      const uint32_t code = (side >> (2 * i)) & 3u;
      double xe = x[kAdvExplicit - 1];
#pragma unroll
      for (int q = kAdvExplicit - 2; q >= 0; q--) xe = next == q ? x[q] : xe;
      const double a3 = code == 1u ? c1[i] : (code == 2u ? -c1[i] : xe);
      next += code == 0u;
      c3[i] = a3 * M.coef3rd;
    }
  } else if (LOL == 0 && MPAS_PAIRED_STENCIL && M.advP) {
    // one access per slot pair
#pragma unroll
    for (int p = 0; p < WA / 2; p++) {
      const mpas_i2v ij = ld_vec<kNtStencil, mpas_i2v>(M.advP, p * E + e);
      const mpas_d2v a1 = ld_vec<kNtStencil, mpas_d2v>(M.advCoefP, p * E + e);
      const mpas_d2v a3 = ld_vec<kNtStencil, mpas_d2v>(M.advCoef3P, p * E + e);
      ic[2 * p] = uint32_t(ij.x);
      ic[2 * p + 1] = uint32_t(ij.y);
      c1[2 * p] = a1.x;
      c1[2 * p + 1] = a1.y;
      c3[2 * p] = a3.x * M.coef3rd;
      c3[2 * p + 1] = a3.y * M.coef3rd;
    }
  } else {
#pragma unroll
    for (int i = 0; i < WA; i++) {
      ic[i] = ldn<LOL == 0 && kNtStencil>(M.adv, i * E + e);
      c1[i] = ldn<LOL == 0 && kNtStencil>(M.advCoef, i * E + e);
      c3[i] = ldn<LOL == 0 && kNtStencil>(M.advCoef3, i * E + e) * M.coef3rd;
    }
  }
#pragma unroll
  for (int i = 0; i < WA; i++) lvc[i] = STG ? int(L) : lvl_cell(M, ic[i]);  // STG: full columns
  // QA advection slots gathered per scheduling phase (QHX, with gathers
  // from global memory): the flux sum takes the slots in order, chunk by
  // chunk, so fewer gathered values are live at once
  constexpr int QA = (!STG && QHX > 0 && QHX < WA) ? (QHX < 2 ? 2 : QHX) : WA;
  for (uint32_t k = k0; k < k1; k += LP) {
    const uint32_t o = e * L + k;
    double xu[LP], he[LP], t1[LP][TPL], t2[LP][TPL];
    int mask[LP];
    ld_lv<LP>(S.xu, o, xu);
    ld_lv<LP>(S.h_edge, o, he);
#pragma unroll
    for (int p = 0; p < LP; p++) mask[p] = M.hoMask ? ldc(M.hoMask, o + p) : 1;
    double ho[LP][TPL];
#pragma unroll
    for (int p = 0; p < LP; p++)
#pragma unroll
      for (int g = 0; g < TPL; g++) ho[p][g] = 0.0;
#pragma unroll
    for (int i0 = 0; i0 < WA; i0 += QA) {
      // STG: the slots' tracers are read from LDS where the flux sum takes
      // them (holding all of them took 152 VGPRs, 3 waves per SIMD)
      double tv[QA][LP][TPL];
      if constexpr (!STG) {
#pragma unroll
        for (int ii = 0; ii < QA && i0 + ii < WA; ii++)
#pragma unroll
          for (int p = 0; p < LP; p++)
            ld_group<TPL>(S.xtr, (ic[i0 + ii] * L + k + p) * NT + t0, tv[ii][p]);
      }
      if (i0 == 0) {
#pragma unroll
        for (int p = 0; p < LP; p++) {
          if (M.advC12) {  // the stencil's first two cells are the edge's cells
#pragma unroll
            for (int g = 0; g < TPL; g++) {
              t1[p][g] = STG ? lds[ic[0] * NT + t0 + g] : tv[0][p][g];
              t2[p][g] = STG ? lds[ic[1] * NT + t0 + g] : tv[1][p][g];
            }
          } else {
            ld_group<TPL>(S.xtr, (uint32_t(cc.x) * L + k + p) * NT + t0, t1[p]);
            ld_group<TPL>(S.xtr, (uint32_t(cc.y) * L + k + p) * NT + t0, t2[p]);
          }
        }
      }
#pragma unroll
      for (int p = 0; p < LP; p++) {
        const int kp = int(k) + p;
        const double F = xu[p] * he[p];
        const double wgt = F * double(mask[p]);
        const double sgn = (F < 0.0) ? -1.0 : double(F > 0.);
#pragma unroll
        for (int ii = 0; ii < QA && i0 + ii < WA; ii++) {
          const int i = i0 + ii;
          const double weightFactor = wgt * (c1[i] + c3[i] * sgn);
          const bool live = i < na && kp < lvc[i];
          double tvi[TPL];
#pragma unroll
          for (int g = 0; g < TPL; g++) tvi[g] = STG ? lds[ic[i] * NT + t0 + g] : tv[ii][p][g];
#pragma unroll
          for (int g = 0; g < TPL; g++) {
            const double x = ho[p][g] + tvi[g] * weightFactor;
            ho[p][g] = live ? x : ho[p][g];
          }
        }
      }
      if (i0 + QA < WA) MPAS_PHASE();
    }
#pragma unroll
    for (int p = 0; p < LP; p++) {
      const int kp = int(k) + p;
      const double F = xu[p] * he[p];
      if (kp < lve) {
        const double tracerWeight = double((mask[p] + 1) & 1) * (dv * 0.5) * F;
        const double ntf1 = smax(0.0, F), ntf2 = smin(0.0, F);
#pragma unroll
        for (int g = 0; g < TPL; g++) {
          const double lo = dv * (ntf1 * t1[p][g] + ntf2 * t2[p][g]);
          ho[p][g] += tracerWeight * (t1[p][g] + t2[p][g]);
          ho[p][g] -= lo;
        }
      }
    }
#pragma unroll
    for (int p = 0; p < LP; p++) st_group<TPL>(S.anti, (o + p) * NT + t0, ho[p]);
  }
}

// compute_substep / accumulate_update / finalize_timestep for one tracer of
// one cell (solve.cc:297-305, 332-340, 350-357)
template <bool NT>
__device__ __forceinline__ void tracer_update(const StageIO &S, uint32_t ol, uint32_t ot,
                                              double te, double hy) {
  if (S.stage == 0) {
    const double h0 = ldo<NT>(S.h0, ol), t0 = ldo<NT>(S.tr0, ot);
    st(S.ytr, ot, (h0 * t0 + S.w_sub * te) / hy);
    stn<NT>(S.ntr, ot, t0 * h0 + S.w_rk * te);
  } else if (S.stage < 3) {
    // every load before the first store (the stores could alias them)
    const double h0 = ldo<NT>(S.h0, ol), t0 = ldo<NT>(S.tr0, ot), q = ldo<NT>(S.ntr, ot);
    st(S.ytr, ot, (h0 * t0 + S.w_sub * te) / hy);
    stn<NT>(S.ntr, ot, q + S.w_rk * te);
  } else {
    st(S.ytr, ot, (ldo<NT>(S.ntr, ot) + S.w_rk * te) / hy);
  }
}

// step_tracers_forward and update_tracers of the tracer-only (ocean) run for
// one tracer of one cell (ocean/solve.cc:368-414, run.cc:10-50): next =
// curr + w_rk tend at stage 0 (copy_tracers, then +=), next += w_rk tend,
// prev = curr + w_sub tend; the last stage writes curr = next + w_rk tend
// (its prev, curr + 0 tend, is not read again).
template <bool NT>
__device__ __forceinline__ void ocean_update(const StageIO &S, uint32_t ot, double te) {
  if (S.stage == 0) {
    const double t0 = ldo<NT>(S.tr0, ot);
    stn<NT>(S.ntr, ot, t0 + S.w_rk * te);
    st(S.ytr, ot, t0 + S.w_sub * te);
  } else if (S.stage < 3) {
    stn<NT>(S.ntr, ot, ldo<NT>(S.ntr, ot) + S.w_rk * te);
    st(S.ytr, ot, ldo<NT>(S.tr0, ot) + S.w_sub * te);
  } else {
    st(S.ytr, ot, ldo<NT>(S.ntr, ot) + S.w_rk * te);
  }
}

// ---- cells (x tracers) -------------------------------------------------------
// Thickness tendency and update, ke and divergence of the new state (the
// cell's first lane stores them), and per tracer either the centred
// tendency with its update or the FCT bounds, upwind tendency and limiter
// factors.  NPL lanes per cell, TPL tracers each (one lane without tracers).
//
// STG (one-level launches of kCellsBlock lanes, DMesh::c2eLoc set): the
// workgroup first copies the edge values its cells read (new and
// provisional u, h_edge, dvEdge, dcEdge, antidiffusive fluxes) of the
// distinct edges of its cells into LDS, ascending edge id, then every cell
// reads them there instead of gathering them from L2; `lds` holds
// 5 + NT arrays of M.stgMax doubles.
//
// CST (one-level FCT launches of kCellsBlock lanes, DMesh::c2cLocC set):
// the workgroup first copies the tracers and thickness of its own cells
// and of the cells across their edges into LDS (`lds`: M.cStgMax records
// of NT tracers, then M.cStgMax thicknesses); every cell reads its own and
// its neighbours' values there by position, and forms the thickness at its
// edges, 0.5 (h[c1] + h[c2]) (solve.cc:31-39, the diagnostics that made
// h_edge), from them instead of loading h_edge.
template <int NT, int W, int LOL = 0, int TPLX = 0, int SCH = -1, int FIX = 0, bool STG = false,
          bool CST = false>
__device__ __forceinline__ void stage_cells(const DMesh &M, const StageIO &S, int item,
                                            double *lds = nullptr) {
  constexpr int TPL = NT > 0 ? (TPLX ? TPLX : tracers_per_lane<NT>()) : 1;
  constexpr int NPL = NT > 0 ? NT / TPL : 1;
  static_assert(!STG || (LOL == 0 && NT > 0 && SCH == MPAS_TRACER_FCT),
                "LDS staging serves one-level FCT launches");
  static_assert(!CST || (LOL == 0 && NT > 0 && NT % 2 == 0 && SCH == MPAS_TRACER_FCT && W <= 6 &&
                         !STG && TPL % 2 == 0),
                "cell staging serves one-level FCT launches of tracer pairs, six edges per cell");
  const uint32_t cmx = CST ? uint32_t(M.cStgMax) : 0u;
  double *const c_tr = lds, *const c_h = lds + NT * cmx;
  int cb = 0;  // CST: the block's first cell
  if constexpr (CST) {
    // every lane of the workgroup takes part, before any returns
    const int lb = item / kCellsBlock;
    cb = lb * M.cStgCpb;
    const int nb = (M.nC + M.cStgCpb - 1) / M.cStgCpb;
    if (lb < nb) {
      const int b0 = ld(M.cStgOff, uint32_t(lb)), nc = ld(M.cStgOff, uint32_t(lb + 1)) - b0;
      // pieces: NT / 2 16-byte pieces of a cell's tracer record, then its
      // thickness; consecutive lanes take consecutive pieces
      constexpr int P = NT / 2 + 1;
      const double2 *tr2 = reinterpret_cast<const double2 *>(S.xtr);
      double2 *ltr = reinterpret_cast<double2 *>(c_tr);
      for (int q = int(threadIdx.x); q < nc * P; q += int(blockDim.x)) {
        const int i = q / P, r = q - i * P;
        const uint32_t x = uint32_t(ld(M.cStgCell, uint32_t(b0 + i)));
        if (r < NT / 2)
          ltr[i * (NT / 2) + r] = ld(tr2, x * uint32_t(NT / 2) + uint32_t(r));
        else
          c_h[i] = ld(S.xh, x);
      }
    }
    __syncthreads();
  }
  const uint32_t mx = STG ? uint32_t(M.stgMax) : 0u;
  double *const s_yu = lds, *const s_xu = lds + mx, *const s_he = lds + 2 * mx;
  double *const s_dv = lds + 3 * mx, *const s_dc = lds + 4 * mx, *const s_an = lds + 5 * mx;
  if constexpr (STG) {
    // every lane of the workgroup takes part, before any returns
    const int lb = item / kCellsBlock;
    const int nb = (M.nC + M.stgCpb - 1) / M.stgCpb;
    if (lb < nb) {
      const int b0 = ld(M.stgOff, uint32_t(lb)), nE = ld(M.stgOff, uint32_t(lb + 1)) - b0;
      for (int i = int(threadIdx.x); i < nE; i += int(blockDim.x)) {
        const uint32_t e = uint32_t(ld(M.stgEdge, uint32_t(b0 + i)));
        if (!S.ocean) s_yu[i] = ld(S.yu, e);
        s_xu[i] = ld(S.xu, e);
        s_he[i] = ld(S.h_edge, e);
        s_dv[i] = ldc(M.dv, e);
        s_dc[i] = ldc(M.dc, e);
        double a[NT];
        ld_group<NT>(S.anti, e * NT, a);
#pragma unroll
        for (int t = 0; t < NT; t++) s_an[uint32_t(i) * NT + t] = a[t];
      }
    }
    __syncthreads();
  }
  int sub;
  uint32_t k0, k1;
  const int c = split_levels<LOL, NPL, FIX>(S, M.nL, item, sub, k0, k1);
  const int t0 = sub * TPL;
  if (c >= M.nC) return;
  const uint32_t L = M.nL, C = M.nC;
  // CST: positions of the cells across the edge slots and (cpos[W]) of the
  // cell itself in the workgroup's cell list
  uint32_t cpos[W + 1];
  if constexpr (CST) {
    const uint64_t loc = ldc(M.c2cLocC, uint32_t(c));
#pragma unroll
    for (int j = 0; j < W; j++) cpos[j] = uint32_t(loc >> (10 * j)) & 1023u;
    cpos[W] = uint32_t(c - cb);
  }
  // CST: h_edge formed from the staged thickness (not in the tracer-only
  // run, whose h_edge is its own fixed field)
  const bool he_h = CST && !S.ocean;
  uint32_t pos[W];  // STG: each slot's position in the workgroup's edge list
  if constexpr (STG) {
    const uint64_t loc = ldc(M.c2eLoc, uint32_t(c));
#pragma unroll
    for (int j = 0; j < W; j++) pos[j] = uint32_t(loc >> (10 * j)) & 1023u;
  }
  const int n = ldc(M.nEoC, c);
  const uint8_t neg = ldc(M.sgnNeg, c);
  const uint32_t rank = ldc(M.ascRank, c);
  const double ar = ldc(M.area, c);
  const bool lead = t0 == 0;
  uint32_t e[W];
  double dv[W];
  // with levels on lanes lane r of the cell's block loads slot r's rows
  const bool shared = LOL == 1 && share_rows<W>(S.nchunk * NPL);
  int base = 0, js = 0;
  uint32_t ej = 0;  // this lane's slot of the row (no run-time index into e[])
  // one level: the cell's NPL tracer lanes (an aligned group of a wave when
  // NPL divides 64) load slots sub, sub + NPL, ... and share them likewise
  constexpr bool GRP = LOL == 0 && NPL > 1 && 64 % NPL == 0 && MPAS_SHARE_ROWS;
  constexpr int MS = (W + NPL - 1) / NPL;
  uint32_t me[MS];
  // shared rows the level loop reads (dcEdge for ke, the other cells of the
  // edges for the tracers), broadcast here: inside the loop the permutes
  // would run under divergent control (only the lead lane forms ke, and the
  // last level chunk may take fewer passes), where an inactive source lane
  // provides no value
  double dcS[W];
  uint32_t opS[W];
  if constexpr (STG) {
    if constexpr (GRP) base = int(threadIdx.x & 63u) - sub;  // for the cOpp rows
#pragma unroll
    for (int j = 0; j < W; j++) {
      e[j] = 0;  // edge ids are not needed: the values come from LDS
      dv[j] = s_dv[pos[j]];
    }
  } else if constexpr (GRP) {
    base = int(threadIdx.x & 63u) - sub;
    double dm[MS];
#pragma unroll
    for (int m = 0; m < MS; m++) {
      me[m] = ldn<kNtCells>(M.c2e, min(sub + m * NPL, W - 1) * C + c);
      dm[m] = ldc(M.dv, me[m]);
    }
    grp_row<NPL>(e, me, base);
    grp_row<NPL>(dv, dm, base);
  } else if (shared) {
    js = row_slot(item, S.nchunk * NPL, W, base);
    ej = ld(M.c2e, js * C + c);
    bcast_row<W, uint32_t>(e, ej, base);
    bcast_row<W, double>(dv, ldc(M.dv, ej), base);
    bcast_row<W, double>(dcS, ldc(M.dc, ej), base);
    if constexpr (NT > 0) bcast_row<W, uint32_t>(opS, ld(M.cOpp, js * C + c), base);
  } else {
#pragma unroll
    for (int j = 0; j < W; j++) e[j] = ldn<LOL == 0 && kNtCells>(M.c2e, j * C + c);
#pragma unroll
    for (int j = 0; j < W; j++) dv[j] = ldc(M.dv, e[j]);
  }
  int lvE[W];  // maxLevelEdgeTop of the cell's edges (partial columns only)
#pragma unroll
  for (int j = 0; j < W; j++) lvE[j] = (!STG && M.mlet) ? ldc(M.mlet, e[j]) : int(L);

  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t ol = c * L + k;
    // the non-tracer outputs are stored after the tracer part's loads: the
    // stores could alias them and would hold the loads back
    double ke_v = 0.0, dvg_v = 0.0, nh_v = 0.0;
    const bool w_ke = lead && !S.ocean;
    bool w_nh = false;
    double yu[W], dc[W];  // new u and dcEdge of the edges, for ke and divergence
    constexpr bool SHKE = GRP && !STG && MPAS_CELLS_SHARE_KE;
    if constexpr (SHKE) {
      // every lane of the cell loads its share of the rows the lead lane
      // needs (the branch on S.ocean is uniform: no lane misses a permute)
      if (!S.ocean) {
        double ym[MS], cm[MS];
#pragma unroll
        for (int m = 0; m < MS; m++) {
          ym[m] = ld(S.yu, me[m] * L + k);
          cm[m] = ldc(M.dc, me[m]);
        }
        grp_row<NPL>(yu, ym, base);
        grp_row<NPL>(dc, cm, base);
      }
    }
    if (lead && !S.ocean) {
      // ke (solve.cc:63-77) and divergence of the new state
      if constexpr (STG) {
#pragma unroll
        for (int j = 0; j < W; j++) {
          yu[j] = s_yu[pos[j]];
          dc[j] = s_dc[pos[j]];
        }
      } else if constexpr (!SHKE) {
#pragma unroll
        for (int j = 0; j < W; j++) yu[j] = ld(S.yu, e[j] * L + k);
        if (shared) {
#pragma unroll
          for (int j = 0; j < W; j++) dc[j] = dcS[j];
        } else {
#pragma unroll
          for (int j = 0; j < W; j++) dc[j] = ldc(M.dc, e[j]);
        }
      }
      double ke = 0.0, dvg = 0.0;
#pragma unroll
      for (int j = 0; j < W; j++) {
        const double kx = ke + 0.25 * dc[j] * dv[j] * (yu[j] * yu[j]);
        const double dx = dvg + (-sign_on_cell(neg, j)) * dv[j] * yu[j];
        ke = (j < n) ? kx : ke;
        dvg = (j < n) ? dx : dvg;
      }
      ke_v = ke / ar;
      dvg_v = dvg / ar;
    }
    double xu[W], he[W];
    if constexpr (STG) {
#pragma unroll
      for (int j = 0; j < W; j++) {
        xu[j] = s_xu[pos[j]];
        he[j] = s_he[pos[j]];
      }
    } else if constexpr (GRP) {
      double xm[MS], hm[MS];
#pragma unroll
      for (int m = 0; m < MS; m++) {
        xm[m] = ld(S.xu, me[m] * L + k);
        if (!he_h) hm[m] = ld(S.h_edge, me[m] * L + k);
      }
      grp_row<NPL>(xu, xm, base);
      if (he_h) {  // uniform: every lane of the cell takes this branch
#pragma unroll
        for (int j = 0; j < W; j++) he[j] = 0.5 * (c_h[cpos[W]] + c_h[cpos[j]]);
      } else {
        grp_row<NPL>(he, hm, base);
      }
    } else {
#pragma unroll
      for (int j = 0; j < W; j++) {
        xu[j] = ld(S.xu, e[j] * L + k);
        he[j] = ld(S.h_edge, e[j] * L + k);
      }
    }
    // compute_tend, thickness part: edges in ascending id (solve.cc:191-205);
    // the tracer-only (ocean) run keeps the thickness fixed
    double hy = 0.0;
    if (!S.ocean) {
      double th;
      {
        double fl[W];
#pragma unroll
        for (int j = 0; j < W; j++) fl[j] = xu[j] * dv[j] * he[j];
        th = sum_ascending<W>(0.0, fl, rank, neg, n);
      }
      th = th / ar;
      constexpr bool NTS = LOL == 0 && kNtNext;
      if (S.stage == 0) {
        const double h0 = ldo<NTS>(S.h0, ol);
        hy = h0 + S.w_sub * th;
        nh_v = h0 + S.w_rk * th;
        w_nh = lead;
      } else if (S.stage < 3) {
        hy = ldo<NTS>(S.h0, ol) + S.w_sub * th;
        nh_v = ldo<NTS>(S.nh, ol) + S.w_rk * th;
        w_nh = lead;
      } else {
        hy = ldo<NTS>(S.nh, ol) + S.w_rk * th;
      }
    }

    if constexpr (NT > 0) {
      // the tracer scheme: SCH when the launch fixes it (each scheme then
      // gets its own registers), else the descriptor's
      const int scheme = SCH >= 0 ? SCH : S.scheme;
      const double invArea = 1.0 / ar;
      const uint32_t ot = ol * NT + t0;
      double tc[TPL];
      double tn[W][TPL];
      if constexpr (CST) {
        // the cell's tracers from LDS (its neighbours' are read there where
        // the FCT sums take them)
#pragma unroll
        for (int g = 0; g < TPL; g++) tc[g] = c_tr[cpos[W] * NT + t0 + g];
      } else {
        ld_group<TPL>(S.xtr, ot, tc);
        uint32_t op[W];
        if constexpr (GRP) {
          uint32_t om[MS];
#pragma unroll
          for (int m = 0; m < MS; m++)
            om[m] = ldn<kNtCells>(M.cOpp, min(sub + m * NPL, W - 1) * C + c);
          grp_row<NPL>(op, om, base);
        } else if (shared) {
#pragma unroll
          for (int j = 0; j < W; j++) op[j] = opS[j];
        } else {
#pragma unroll
          for (int j = 0; j < W; j++) op[j] = ldn<LOL == 0 && kNtCells>(M.cOpp, j * C + c);
        }
#pragma unroll
        for (int j = 0; j < W; j++) ld_group<TPL>(S.xtr, (op[j] * L + k) * NT + t0, tn[j]);
      }
      if (scheme == MPAS_TRACER_CENTRED) {
        // compute_scalar_tend (solve.cc:237-270), ascending edge id, then
        // tracer_hmix_del2_tend (ocean/solve.cc:328-366), file order
#pragma unroll
        for (int g = 0; g < TPL; g++) {
          double term[W];
#pragma unroll
          for (int j = 0; j < W; j++) {
            const bool first = (neg >> j) & 1;
            const double tracer_edge =
                first ? 0.5 * (tc[g] + tn[j][g]) : 0.5 * (tn[j][g] + tc[g]);
            term[j] = (xu[j] * dv[j] * he[j] * tracer_edge) / ar;
          }
          double te = sum_ascending<W>(0.0, term, rank, neg, n);
          if (S.use_hmix) {
#pragma unroll
            for (int j = 0; j < W; j++) {
              const bool first = (neg >> j) & 1;
              const double rho_scaling = M.msd * S.kappa * dv[j] / ldc(M.dc, e[j]);
              const double turb = first ? tn[j][g] - tc[g] : tc[g] - tn[j][g];
              const double flux = he[j] * turb * rho_scaling;
              const double x = te - (first ? -1.0 : 1.0) * flux * invArea;
              // levels below the edge's maxLevelEdgeTop only (ocean/solve.cc:344-363)
              te = (j < n && int(k) < lvE[j]) ? x : te;
            }
          }
          tracer_update<LOL == 0 && kNtNext>(S, ol, ot + g, te, hy);
        }
      } else if (scheme == MPAS_TRACER_FCT) {
        // calc_provisional_layer_thicknesses (ocean/solve.cc:74-96)
        const double dtInvAreaCell = S.dt / ar;
        const double lt = CST ? c_h[cpos[W]] : ld(S.xh, ol);
        double hp = lt;
#pragma unroll
        for (int j = 0; j < W; j++) {
          const double signedFactor = dtInvAreaCell * dv[j] * sign_on_cell(neg, j);
          const double x = hp + signedFactor * (xu[j] * he[j]);
          hp = (j < n) ? x : hp;
        }
        const double hProvInv = 1.0 / hp;
        double an[W][TPL];
        if constexpr (STG) {
#pragma unroll
          for (int j = 0; j < W; j++)
#pragma unroll
            for (int g = 0; g < TPL; g++) an[j][g] = s_an[pos[j] * NT + t0 + g];
        } else {
#pragma unroll
        for (int j = 0; j < W; j++) ld_group<TPL>(S.anti, (e[j] * L + k) * NT + t0, an[j]);
        }
        double rf[2 * TPL], tl[TPL];
#pragma unroll
        for (int g = 0; g < TPL; g++) {
          // bounds over cellsOnCell (195-214); upwind tendency and
          // antidiffusive in/out sums, file order (232-249)
          double tmax = tc[g], tmin = tc[g], te = 0.0, fin = 0.0, fout = 0.0;
#pragma unroll
          for (int j = 0; j < W; j++) {
            const bool live = j < n;
            const double tnj = CST ? c_tr[cpos[j] * NT + t0 + g] : tn[j][g];
            tmax = live ? smax(tmax, tnj) : tmax;
            tmin = live ? smin(tmin, tnj) : tmin;
            const double signedFactor = sign_on_cell(neg, j) * invArea;
            const double F = xu[j] * he[j];
            const double ntf1 = smax(0.0, F), ntf2 = smin(0.0, F);
            const bool first = (neg >> j) & 1;
            const double t1 = first ? tc[g] : tnj;
            const double t2 = first ? tnj : tc[g];
            const double lo = dv[j] * (ntf1 * t1 + ntf2 * t2);
            const double xe = te + signedFactor * lo;
            const double xi = fin + smax(0.0, signedFactor * an[j][g]);
            const double xo = fout + smin(0.0, signedFactor * an[j][g]);
            te = live ? xe : te;
            fin = live ? xi : fin;
            fout = live ? xo : fout;
          }
          // limiter factors (253-266), stored (in, out) per tracer
          const double eps = 1.e-10;
          const double tracerUpwindNew = (tc[g] * lt + S.dt * te) * hProvInv;
          const double tracerMaxNew = tracerUpwindNew + S.dt * fin * hProvInv;
          const double tracerMinNew = tracerUpwindNew + S.dt * fout * hProvInv;
          double scaleFactor =
              (tmax - tracerUpwindNew) / (tracerMaxNew - tracerUpwindNew + eps);
          rf[2 * g] = smin(1.0, smax(0.0, scaleFactor));
          scaleFactor = (tracerUpwindNew - tmin) / (tracerUpwindNew - tracerMinNew + eps);
          rf[2 * g + 1] = smin(1.0, smax(0.0, scaleFactor));
          tl[g] = te;
        }
        if (!S.keep_rfac) {
#pragma unroll
          for (int g = 0; g < TPL; g++) {
            const double pair[2] = {rf[2 * g], rf[2 * g + 1]};
            st_group<2>(S.rfac, (ot + g) * 2, pair);
          }
        }
        st_group<TPL>(S.tlo, ot, tl);
      }
    }
    if (w_ke) {
      st(S.ke_y, ol, ke_v);
      if (S.use_visc) st(S.div_y, ol, dvg_v);
    }
    if (w_nh) stn<LOL == 0 && kNtNext>(S.nh, ol, nh_v);
    if (lead && !S.ocean) st(S.yh, ol, hy);
  }
}

// ---- cells x tracers: FCT rescale + accumulate, del2 mixing, update --------
//
// STG (one-level launches of kTracerBlock lanes, DMesh::c2cLoc set): the
// workgroup first copies the limiter factors and tracers of its own cells
// and of the cells across their edges into LDS (`lds`: M.tStgMax records
// of 2 NT factors, then M.tStgMax records of NT tracers), then every cell
// reads its own and its neighbours' values there by position, where the
// sums take them.  (Forming h_edge from staged thickness as well, which
// the copy would add as one scattered load per cell, made the kernel
// slower: 82.5 against 66.0 us per launch at level 8.)  A cell's neighbours are mostly the block's own cells or
// shared with other cells of the block, so the copy reads each record once
// where the gathers read it about four times.
template <int NT, int W, int LOL = 0, int TPLX = 0, int FIX = 0, bool STG = false>
__device__ __forceinline__ void stage_tracer(const DMesh &M, const StageIO &S, int item,
                                             double *lds = nullptr) {
  constexpr int TPL = TPLX ? TPLX : tracers_per_lane<NT>(), NPL = NT / TPL;
  static_assert(!STG || (LOL == 0 && W <= 6 && NT % 2 == 0),
                "LDS staging serves one-level launches, at most six edges per cell, even tracer counts");
  const uint32_t mx = STG ? uint32_t(M.tStgMax) : 0u;
  double *const s_rf = lds, *const s_tr = lds + 2 * NT * mx;
  int cb = 0;  // STG: the block's first cell
  if constexpr (STG) {
    // every lane of the workgroup takes part, before any returns
    const int lb = item / kTracerBlock;
    cb = lb * M.tStgCpb;
    const int nb = (M.nC + M.tStgCpb - 1) / M.tStgCpb;
    if (lb < nb) {
      const int b0 = ld(M.tStgOff, uint32_t(lb)), nc = ld(M.tStgOff, uint32_t(lb + 1)) - b0;
      // 16-byte pieces: NT of a cell's factor record (2 NT doubles), then
      // NT / 2 of its tracer record; consecutive lanes take consecutive
      // pieces, so the block's own cells are copied in one contiguous run
      constexpr int P = NT + NT / 2;
      const double2 *rf2 = reinterpret_cast<const double2 *>(S.rfac);
      const double2 *tr2 = reinterpret_cast<const double2 *>(S.xtr);
      double2 *lrf = reinterpret_cast<double2 *>(s_rf);
      double2 *ltr = reinterpret_cast<double2 *>(s_tr);
      for (int q = int(threadIdx.x); q < nc * P; q += int(blockDim.x)) {
        const int i = q / P, r = q - i * P;
        const uint32_t x = uint32_t(ld(M.tStgCell, uint32_t(b0 + i)));
        if (r < NT)
          lrf[i * NT + r] = ld(rf2, x * uint32_t(NT) + uint32_t(r));
        else
          ltr[i * (NT / 2) + (r - NT)] = ld(tr2, x * uint32_t(NT / 2) + uint32_t(r - NT));
      }
    }
    __syncthreads();
  }
  int sub;
  uint32_t k0, k1;
  const int c = split_levels<LOL, NPL, FIX>(S, M.nL, item, sub, k0, k1);
  const int t0 = sub * TPL;
  if (c >= M.nC) return;
  const uint32_t L = M.nL, C = M.nC;
  const int n = ldc(M.nEoC, c);
  const uint8_t neg = ldc(M.sgnNeg, c);
  const double invAreaCell1 = 1.0 / ldc(M.area, c);
  uint32_t e[W], op[W];
  // STG: positions of the cell (pos[W]) and of its neighbours in the list
  uint32_t pos[W + 1];
  if constexpr (STG) {
    const uint64_t loc = ldc(M.c2cLoc, uint32_t(c));
#pragma unroll
    for (int j = 0; j < W; j++) pos[j] = uint32_t(loc >> (10 * j)) & 1023u;
    pos[W] = uint32_t(c - cb);
  }
  // with levels on lanes lane r of the cell's block loads slot r's rows
  const bool shared = LOL == 1 && share_rows<W>(S.nchunk * NPL);
  int base = 0, js = 0;
  uint32_t ej = 0;  // this lane's slot of the row (no run-time index into e[])
  // one level: the cell's NPL tracer lanes (an aligned group of a wave when
  // NPL divides 64) load slots sub, sub + NPL, ... and share them likewise
  constexpr bool GRP = LOL == 0 && NPL > 1 && 64 % NPL == 0 && MPAS_SHARE_ROWS;
  constexpr int MS = (W + NPL - 1) / NPL;
  uint32_t me[MS];
  // the mixing coefficients of the edges, shared rows broadcast before the
  // level loop (whose pass count may differ between the level chunks)
  double rhoS[W];
  if constexpr (GRP) {
    base = int(threadIdx.x & 63u) - sub;
    uint32_t mo[MS];
#pragma unroll
    for (int m = 0; m < MS; m++) {
      const int j = min(sub + m * NPL, W - 1);
      me[m] = ldn<kNtCells>(M.c2e, j * C + c);
      if constexpr (!STG) mo[m] = ldn<kNtCells>(M.cOpp, j * C + c);
    }
    grp_row<NPL>(e, me, base);
    if constexpr (!STG) grp_row<NPL>(op, mo, base);
  } else if (shared) {
    js = row_slot(item, S.nchunk * NPL, W, base);
    ej = ld(M.c2e, js * C + c);
    bcast_row<W, uint32_t>(e, ej, base);
    bcast_row<W, uint32_t>(op, ld(M.cOpp, js * C + c), base);
    if (S.use_hmix)
      bcast_row<W, double>(rhoS, M.msd * S.kappa * ldc(M.dv, ej) / ldc(M.dc, ej), base);
  } else {
#pragma unroll
    for (int j = 0; j < W; j++) {
      e[j] = ldn<LOL == 0 && kNtCells>(M.c2e, j * C + c);
      if constexpr (!STG) op[j] = ldn<LOL == 0 && kNtCells>(M.cOpp, j * C + c);
    }
  }
  MPAS_PHASE();
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t ol = c * L + k, ot = ol * NT + t0;
    double te[TPL], rs[2 * TPL];
    ld_group<TPL>(S.tlo, ot, te);
    if constexpr (STG) {
#pragma unroll
      for (int g = 0; g < 2 * TPL; g++) rs[g] = s_rf[pos[W] * (2 * NT) + 2 * t0 + g];
    } else {
#pragma unroll
      for (int g = 0; g < TPL; g++) {
        double pair[2];
        ld_group<2>(S.rfac, (ot + g) * 2, pair);
        rs[2 * g] = pair[0];
        rs[2 * g + 1] = pair[1];
      }
    }
    double an[W][TPL], ro[W][2 * TPL];
#pragma unroll
    for (int j = 0; j < W; j++) {
      ld_group<TPL>(S.anti, (e[j] * L + k) * NT + t0, an[j]);
      if constexpr (!STG) {
        const uint32_t oo = (op[j] * L + k) * NT + t0;
#pragma unroll
        for (int g = 0; g < TPL; g++) {
          double pair[2];
          ld_group<2>(S.rfac, (oo + g) * 2, pair);
          ro[j][2 * g] = pair[0];
          ro[j][2 * g + 1] = pair[1];
        }
      }
    }
    MPAS_PHASE();
    // rescale (ocean/solve.cc:270-282) and accumulate (301-316), file order
#pragma unroll
    for (int j = 0; j < W; j++) {
      const bool first = (neg >> j) & 1;  // c == m.cells(e)[0]
      const double signedFactor = (first ? -1.0 : 1.0) * invAreaCell1;
#pragma unroll
      for (int g = 0; g < TPL; g++) {
        const double rin = rs[2 * g], rout = rs[2 * g + 1];
        // STG: the neighbour's factors read from LDS here
        const double oin = STG ? s_rf[pos[j] * (2 * NT) + 2 * (t0 + g)] : ro[j][2 * g];
        const double oout = STG ? s_rf[pos[j] * (2 * NT) + 2 * (t0 + g) + 1] : ro[j][2 * g + 1];
        const double in1 = first ? rin : oin, out1 = first ? rout : oout;
        const double in2 = first ? oin : rin, out2 = first ? oout : rout;
        const double x = an[j][g];
        const double hof = smax(0.0, x) * smin(out1, in2) + smin(0.0, x) * smin(in1, out2);
        const double y = te[g] + signedFactor * hof;
        te[g] = (j < n) ? y : te[g];
      }
    }
    if (S.use_hmix) {
      // tracer_hmix_del2_tend (ocean/solve.cc:328-366)
      double tc[TPL], tn[W][TPL], he[W], rho[W];
      if constexpr (STG) {
#pragma unroll
        for (int g = 0; g < TPL; g++) tc[g] = s_tr[pos[W] * NT + t0 + g];
      } else {
        ld_group<TPL>(S.xtr, ot, tc);
      }
      if constexpr (GRP) {
        double hm[MS], rm[MS];
#pragma unroll
        for (int m = 0; m < MS; m++) {
          hm[m] = ld(S.h_edge, me[m] * L + k);
          rm[m] = M.msd * S.kappa * ldc(M.dv, me[m]) / ldc(M.dc, me[m]);
        }
        if constexpr (!STG) {
#pragma unroll
          for (int j = 0; j < W; j++) ld_group<TPL>(S.xtr, (op[j] * L + k) * NT + t0, tn[j]);
        }
        grp_row<NPL>(he, hm, base);
        grp_row<NPL>(rho, rm, base);
      } else {
#pragma unroll
        for (int j = 0; j < W; j++) {
          if constexpr (!STG) ld_group<TPL>(S.xtr, (op[j] * L + k) * NT + t0, tn[j]);
          he[j] = ld(S.h_edge, e[j] * L + k);
        }
      }
      if constexpr (!GRP) {
        if (shared) {
#pragma unroll
          for (int j = 0; j < W; j++) rho[j] = rhoS[j];
        } else {
#pragma unroll
          for (int j = 0; j < W; j++)
            rho[j] = M.msd * S.kappa * ldc(M.dv, e[j]) / ldc(M.dc, e[j]);
        }
      }
      MPAS_PHASE();
#pragma unroll
      for (int j = 0; j < W; j++) {
        const bool first = (neg >> j) & 1;
#pragma unroll
        for (int g = 0; g < TPL; g++) {
          const double tnj = STG ? s_tr[pos[j] * NT + t0 + g] : tn[j][g];
          const double turb = first ? tnj - tc[g] : tc[g] - tnj;
          const double flux = he[j] * turb * rho[j];
          const double y = te[g] - (first ? -1.0 : 1.0) * flux * invAreaCell1;
          // every level is below maxLevelEdgeTop here: the fused FCT runs
          // refuse partial columns (ocean/solve.cc:344-363 stops there)
          te[g] = (j < n) ? y : te[g];
        }
      }
    }
    if (S.ocean) {
#pragma unroll
      for (int g = 0; g < TPL; g++) ocean_update<LOL == 0 && kNtNext>(S, ot + g, te[g]);
    } else {
      const double hy = ld(S.yh, ol);
#pragma unroll
      for (int g = 0; g < TPL; g++) tracer_update<LOL == 0 && kNtNext>(S, ol, ot + g, te[g], hy);
    }
  }
}

}  // namespace mpasx
