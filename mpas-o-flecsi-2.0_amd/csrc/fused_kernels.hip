// Stage-fused RK4 step (apps/sw/run.cc:13-70 with the FCT tracer tasks of
// apps/ocean/run.cc:19-42 as the tracer tendency).
//
// One RK stage of the reference is 6 to 14 FleCSI task launches; here it is
//   k_stage_edges  (edges)    tend_u (+ del2 viscosity), u substep/accumulate
//   k_stage_ho     (edges)    FCT: normal thickness flux, high-order minus
//                             low-order (antidiffusive) flux
//   k_stage_cells  (cells)    tend_h, h substep/accumulate, ke and divergence
//                             of the new state, FCT bounds, upwind tendency and
//                             limiter factors (or the centred tracer tendency)
//   k_stage_tracer (cells)    FCT: rescaled antidiffusive fluxes, del2 mixing,
//                             tracer substep/accumulate
//   k_diag_vertex  (vertices) vorticity and pv_vertex of the new state
//   k_diag_edge    (edges)    h_edge and pv_edge of the new state
// (the bodies are in stage_kernels.h).
// Each value is formed with the operations and in the order of the per-task
// path (the same device operators), so the fused step reproduces the
// reference task sequence bit for bit; the reference's copy tasks
// (init_provis, setup_timestep, timeshift) become buffer roles.
#include "ctx.h"
#include "stage_kernels.h"

namespace mpasx {

namespace {

constexpr int BS = 256;
// Workgroup size of each stage kernel.
#ifndef MPAS_BS_EDGES
#define MPAS_BS_EDGES 256
#endif
#ifndef MPAS_BS_HO
#define MPAS_BS_HO 256
#endif
#ifndef MPAS_BS_CELLS
#define MPAS_BS_CELLS 256
#endif
#ifndef MPAS_BS_TRACER
#define MPAS_BS_TRACER 256
#endif
#ifndef MPAS_BS_DIAG
#define MPAS_BS_DIAG 256
#endif
constexpr int BS_EDGES = MPAS_BS_EDGES, BS_HO = MPAS_BS_HO, BS_CELLS = MPAS_BS_CELLS,
              BS_TRACER = MPAS_BS_TRACER, BS_DIAG = MPAS_BS_DIAG;

// Waves per SIMD each stage kernel is compiled for (0: the compiler's
// choice).  A target caps the registers per lane at 512 / waves.
#ifndef MPAS_WPE_EDGES
#define MPAS_WPE_EDGES 0
#endif
#ifndef MPAS_WPE_HO
#define MPAS_WPE_HO 0
#endif
#ifndef MPAS_WPE_CELLS
#define MPAS_WPE_CELLS 0
#endif
// Resident workgroups per CU of the tracer-update kernel.  Its 68 VGPRs
// would allow 7 waves per SIMD, but its gathers of the neighbours'
// limiter factors ran faster with at most 4 workgroups of 256 lanes per CU
// (83 against 86 us per launch on the level-8 workload, 4 tracers).  A
// dynamic LDS reservation the kernel never touches sets that limit: the
// CU's 160 KiB hold no more than WG such reservations.  0: no limit.
#ifndef MPAS_WG_PER_CU_TRACER
#define MPAS_WG_PER_CU_TRACER 4
#endif
#ifndef MPAS_WG_PER_CU_TRACER_LANES  // the same with levels on lanes
#define MPAS_WG_PER_CU_TRACER_LANES MPAS_WG_PER_CU_TRACER
#endif
// The same for the LDS-staged tracer update, whose neighbour values come
// from LDS: it ran fastest without a limit (66.0 us per launch, against
// 89.8, 73.8 and 67.4 us with 4, 6 and 8 workgroups per CU).
#ifndef MPAS_WG_PER_CU_TRACER_LDS
#define MPAS_WG_PER_CU_TRACER_LDS 0
#endif
constexpr unsigned kLdsPerCu = 160u * 1024u;
constexpr unsigned occupancy_lds(int wg) {
  return wg > 0 ? (kLdsPerCu / unsigned(wg + 1) + kLdsPerCu / unsigned(wg)) / 2 : 0u;
}
static_assert(kLdsPerCu / occupancy_lds(4) == 4 && kLdsPerCu / occupancy_lds(5) == 5 &&
                  kLdsPerCu / occupancy_lds(6) == 6,
              "occupancy_lds picks WG workgroups per CU");
#ifndef MPAS_WPE_TRACER
#define MPAS_WPE_TRACER 0
#endif
#define MPAS_WAVES(n) __attribute__((amdgpu_waves_per_eu((n) > 0 ? (n) : 1, (n) > 0 ? (n) : 8)))

// LOL: 0 for one entity per lane (all its levels), 1 for one level per
// lane with an entity's levels on adjacent lanes (contexts of more than one
// level: the accesses of a wave are then contiguous runs across levels).
template <int WE, int LOL, int LP = 1>
__global__ __launch_bounds__(BS_EDGES) MPAS_WAVES(MPAS_WPE_EDGES) void k_stage_edges(DMesh M,
                                                                               StageIO S) {
  stage_edges<WE, LOL, LP, LP>(M, S, xcd_item());
}
// Tracers per lane of the cell kernels (0: pairs when the count is even;
// the high-order flux kernel's count, MPAS_TPL_HO, and the one-level
// tracer update's, MPAS_TPL_TRACER, are in dmesh.h).
#ifndef MPAS_TPL_CELLS
#define MPAS_TPL_CELLS 0
#endif
// With levels on lanes the tracer update takes pairs: the lanes of a cell
// then cover its levels x tracer pairs, and one lane per pair halves the
// cell's index and neighbour-factor gathers (level 7, 32 levels, 2
// tracers: 346 -> 276 us); one-level launches keep one tracer per lane.
#ifndef MPAS_TPL_TRACER_LANES
#define MPAS_TPL_TRACER_LANES 2
#endif
constexpr int tpl_tracer(int lol) { return lol ? MPAS_TPL_TRACER_LANES : MPAS_TPL_TRACER; }
template <int NT, int X>
constexpr int lanes_per_cell() {
  return NT > 0 ? NT / (X && NT % X == 0 ? X : tracers_per_lane<NT>()) : 1;
}
template <int NT>
constexpr int tpl_of(int x) {
  return x && NT % x == 0 ? x : 0;
}
template <int NT, int WA, int LOL, int LP = 1>
__global__ __launch_bounds__(BS_HO) MPAS_WAVES(MPAS_WPE_HO) void k_stage_ho(DMesh M, StageIO S) {
  stage_ho<NT, WA, LOL, tpl_of<NT>(MPAS_TPL_HO), LP, LP>(M, S, xcd_item());
}
// SCH: the tracer scheme, fixed per instantiation so that the level-8 FCT
// kernel does not carry the centred branch's registers (238 VGPRs with
// both, 117 with FCT alone: twice the waves per SIMD).
template <int NT, int W, int LOL, int SCH>
__global__ __launch_bounds__(BS_CELLS) MPAS_WAVES(MPAS_WPE_CELLS) void k_stage_cells(DMesh M,
                                                                               StageIO S) {
  stage_cells<NT, W, LOL, tpl_of<NT>(MPAS_TPL_CELLS), SCH, 1>(M, S, xcd_item());
}
// One-level momentum with the edgesOnEdge values staged in LDS
// (DMesh::e2eLoc); dynamic LDS of 3 * M.eStgMax doubles.
static_assert(BS_EDGES == kEdgesBlock, "the LDS staging lists are built per kEdgesBlock lanes");
template <int WE>
__global__ __launch_bounds__(BS_EDGES) MPAS_WAVES(MPAS_WPE_EDGES) void k_stage_edges_lds(DMesh M,
                                                                                  StageIO S) {
  extern __shared__ double lds[];
  stage_edges<WE, 0, 1, 0, MPAS_QCH, true>(M, S, xcd_item(), lds);
}
// One-level high-order flux with the advection cells' tracers staged in
// LDS (DMesh::a2cLoc); dynamic LDS of NT * M.aStgMax doubles.
static_assert(BS_HO == kHoBlock, "the LDS staging lists are built per kHoBlock lanes");
template <int NT, int WA>
__global__ __launch_bounds__(BS_HO) MPAS_WAVES(MPAS_WPE_HO) void k_stage_ho_lds(DMesh M,
                                                                         StageIO S) {
  extern __shared__ double lds[];
  stage_ho<NT, WA, 0, tpl_of<NT>(MPAS_TPL_HO), 1, 1, true>(M, S, xcd_item(), lds);
}
// The staged momentum kernel (the first `eblocks` workgroups) and the
// staged high-order flux kernel (the rest) in one launch: the two read
// the same stage input and are independent.
#ifndef MPAS_MERGE_EH
#define MPAS_MERGE_EH 1
#endif
template <int WE, int NT>
__global__ __launch_bounds__(BS_EDGES) void k_stage_edges_ho(DMesh M, StageIO S, int eblocks) {
  extern __shared__ double lds[];
  const int b = int(blockIdx.x);
  if (b < eblocks)
    stage_edges<WE, 0, 1, 0, MPAS_QCH, true>(M, S, xcd_item_in(b, eblocks), lds);
  else
    stage_ho<NT, WE, 0, tpl_of<NT>(MPAS_TPL_HO), 1, 1, true>(
        M, S, xcd_item_in(b - eblocks, int(gridDim.x) - eblocks), lds);
}
// One-level FCT cells with the workgroup's edge values staged in LDS
// (DMesh::c2eLoc); dynamic LDS of (5 + NT) * M.stgMax doubles.
static_assert(BS_CELLS == kCellsBlock, "the LDS staging lists are built per kCellsBlock lanes");
template <int NT, int W>
__global__ __launch_bounds__(BS_CELLS) MPAS_WAVES(MPAS_WPE_CELLS) void k_stage_cells_lds(DMesh M,
                                                                                   StageIO S) {
  extern __shared__ double lds[];
  stage_cells<NT, W, 0, tpl_of<NT>(MPAS_TPL_CELLS), MPAS_TRACER_FCT, 1, true>(M, S, xcd_item(),
                                                                                lds);
}
// One-level FCT cells with its cells' and their neighbours' tracers and
// thickness staged in LDS (DMesh::c2cLocC); dynamic LDS of (NT + 1) *
// M.cStgMax doubles.
template <int NT, int W>
__global__ __launch_bounds__(BS_CELLS) MPAS_WAVES(MPAS_WPE_CELLS) void k_stage_cells_cst(DMesh M,
                                                                                   StageIO S) {
  extern __shared__ double lds[];
  stage_cells<NT, W, 0, tpl_of<NT>(MPAS_TPL_CELLS), MPAS_TRACER_FCT, 1, false, true>(
      M, S, xcd_item(), lds);
}
template <int NT, int W, int LOL>
__global__ __launch_bounds__(BS_TRACER) MPAS_WAVES(MPAS_WPE_TRACER) void k_stage_tracer(DMesh M,
                                                                                 StageIO S) {
  stage_tracer<NT, W, LOL, tpl_of<NT>(tpl_tracer(LOL)), 1>(M, S, xcd_item());
}
// One-level tracer update with the limiter factors and tracers of its
// cells and of their neighbours staged in LDS (DMesh::c2cLoc); dynamic LDS
// of 3 * NT * M.tStgMax doubles, or the occupancy reservation if larger.
static_assert(BS_TRACER == kTracerBlock, "the LDS staging lists are built per kTracerBlock lanes");
template <int NT, int W>
__global__ __launch_bounds__(BS_TRACER) MPAS_WAVES(MPAS_WPE_TRACER) void k_stage_tracer_lds(
    DMesh M, StageIO S) {
  extern __shared__ double lds[];
  stage_tracer<NT, W, 0, tpl_of<NT>(tpl_tracer(0)), 1, true>(M, S, xcd_item(), lds);
}

// The staged tracer update (the first `tblocks` workgroups) and the
// vertex diagnostics of the stage's output (the rest) in one launch: the
// two are independent, and the vertex workgroups fill the CUs the tracer
// update's last workgroups leave idle.
#ifndef MPAS_MERGE_TV
#define MPAS_MERGE_TV 1
#endif
static_assert(BS_TRACER == BS_DIAG, "one workgroup size for both parts");
template <int NT, int W>
__global__ __launch_bounds__(BS_TRACER) MPAS_WAVES(MPAS_WPE_TRACER) void k_stage_tracer_vertex(
    DMesh M, StageIO S, int tblocks, double *vorticity, double *pv_vertex) {
  extern __shared__ double lds[];
  const int b = int(blockIdx.x);
  if (b < tblocks) {
    stage_tracer<NT, W, 0, tpl_of<NT>(tpl_tracer(0)), 1, true>(M, S, xcd_item_in(b, tblocks),
                                                                lds);
  } else {
    const int v = xcd_item_in(b - tblocks, int(gridDim.x) - tblocks);
    if (v < M.nV) diag_vertex<1>(M, v, S.yh, S.yu, nullptr, nullptr, vorticity, pv_vertex);
  }
}
template <int LOL, int LP = 1>
__global__ __launch_bounds__(BS_DIAG) void k_diag_vertex(DMesh M, StageIO S,
                                                         const double *__restrict__ h,
                                                         const double *__restrict__ u,
                                                         double *vorticity, double *pv_vertex) {
  int sub;
  uint32_t k0, k1;
  const int v = split_levels<LOL, 1, LP>(S, M.nL, xcd_item(), sub, k0, k1);
  if (v < M.nV) diag_vertex<LP>(M, v, h, u, nullptr, nullptr, vorticity, pv_vertex, k0, k1);
}

template <int LOL, int LP = 1>
__global__ __launch_bounds__(BS_DIAG) void k_diag_edge(DMesh M, StageIO S,
                                                       const double *__restrict__ h,
                                                       const double *__restrict__ pv_vertex,
                                                       double *h_edge, double *pv_edge) {
  int sub;
  uint32_t k0, k1;
  const int e = split_levels<LOL, 1, LP>(S, M.nL, xcd_item(), sub, k0, k1);
  if (e < M.nE) diag_edge<LP>(M, e, h, pv_vertex, h_edge, pv_edge, k0, k1);
}

// The layout of a context's launches and the StageIO fields it reads.
int layout_of(const DMesh &M) { return M.nL > 1 ? 1 : 0; }
StageIO with_layout(StageIO S, const DMesh &M) {
  S.lpl = 1;
  S.nchunk = M.nL;
  return S;
}

// ke and divergence of a state (the cell part of the initial diagnostics)
__global__ __launch_bounds__(BS) void k_diag_cell(DMesh M, const double *__restrict__ u,
                                                  double *ke, double *div, int use_visc) {
  const int c = xcd_item();
  if (c >= M.nC) return;
  const uint32_t L = M.nL, C = M.nC;
  const int n = ld(M.nEoC, c);
  const uint8_t neg = ld(M.sgnNeg, c);
  const double ar = ld(M.area, c);
  for (uint32_t k = 0; k < L; k++) {
    double s = 0.0, d = 0.0;
    for (int j = 0; j < n; j++) {
      const uint32_t e = ld(M.c2e, j * C + c);
      const double ue = ld(u, e * L + k);
      s += 0.25 * ld(M.dc, e) * ld(M.dv, e) * (ue * ue);
      if (use_visc) d += (-sign_on_cell(neg, j)) * ld(M.dv, e) * ue;
    }
    st(ke, c * L + k, s / ar);
    if (use_visc) st(div, c * L + k, d / ar);
  }
}

// Compulsory HBM bytes of one launch: every array element the launch reads
// or writes, counted once (gathered neighbours included once).  This is
// the algorithmic traffic the roofline of each kernel is measured against.
struct Bytes {
  double edges, ho, cells, tracer, vertex, edge_diag;
};

Bytes stage_bytes(const DMesh &M, const StageIO &S, double sum_eoe, double sum_adv,
                  double sum_eoc) {
  const double C = M.nC, E = M.nE, V = M.nV, L = M.nL, T = M.nT, d = 8.0;
  const bool fct = T > 0 && S.scheme == MPAS_TRACER_FCT;
  const bool cen = T > 0 && S.scheme == MPAS_TRACER_CENTRED;
  // state touches of an RK4 stage: stage 0 reads prev (= provis) and writes
  // provis and next; stages 1-2 read prev, read+write next, write provis;
  // stage 3 reads next and writes the final state
  const double touches = S.stage == 0 ? 2 : S.stage < 3 ? 4 : 2;
  Bytes b{};
  // k_stage_edges: momentum
  if (S.freeze_u) {
    b.edges = S.stage < 3 ? 2 * d * L * E : 0;
  } else {
    b.edges = 8 * E + d * E + 1 * E + 4 * sum_eoe + d * sum_eoe;   // e2c, dc, eoe, woe
    if (M.e2eLoc)  // slot positions and the workgroups' lists instead of the ids
      b.edges += 16 * E + 4.0 * M.eStgLen + 4 * ((E + kEdgesBlock - 1) / kEdgesBlock + 1) -
                 4 * sum_eoe;
    b.edges += d * L * (3 * E + 2 * C) + d * C;                    // xu h_edge pv_edge, ke xh, h_s
    if (S.use_visc) b.edges += 8 * E + d * E + d * L * (C + V);    // e2v, dv, div, vort
    b.edges += touches * d * L * E;                                // u0 / nu / yu
  }
  // k_stage_ho: antidiffusive flux per edge and tracer
  // advection stencil: id + advCoef (+ advCoef_3rd, or in compact form a
  // side code and the explicit values per edge for one-level kernels)
  const bool compact = MPAS_COMPACT_C3 && M.advSide && L == 1;
  const double stencil = compact ? 12 * sum_adv + 4 * E + d * kAdvExplicit * E : 20 * sum_adv;
  if (fct) b.ho = 8 * E + d * E + E + stencil + 2 * d * L * E + d * L * T * (C + E);
  if (fct && M.a2cLoc)  // slot positions and the workgroups' lists instead of the ids
    b.ho += 16 * E + 4.0 * M.aStgLen + 4 * ((E + M.aStgEpb - 1) / M.aStgEpb + 1) - 4 * sum_adv;
  // k_stage_cells
  b.cells = 14 * C + 4 * sum_eoc + 2 * d * E;                      // nEoC sgn rank area, c2e, dv dc
  b.cells += d * L * 3 * E;                                        // yu xu h_edge
  b.cells += d * L * C * (touches + 1 + (S.use_visc ? 1 : 0));     // h0/nh/yh, ke_y, div_y
  if (fct) b.cells += 4 * sum_eoc + d * L * (C + T * C + T * E + 3 * T * C);  // cOpp, xh, xtr, anti, rfac tlo
  if (fct && M.c2cLocC && !M.c2eLoc && L == 1)  // positions and lists instead of cOpp
    b.cells += 8 * C + 4.0 * M.cStgLen + 4 * ((C + M.cStgCpb - 1) / M.cStgCpb + 1) -
               4 * sum_eoc - (S.ocean ? 0 : d * E);  // h_edge formed from the thickness
  if (cen) b.cells += 4 * sum_eoc + d * L * T * C * (touches + 1);  // cOpp, xtr + tr0/ntr/ytr
  // k_stage_tracer
  if (fct) {
    b.tracer = 10 * C + 8 * sum_eoc;                               // nEoC sgn area, c2e cOpp
    if (M.c2cLoc && L == 1)  // slot positions and the workgroups' lists instead of cOpp
      b.tracer += 8 * C + 4.0 * M.tStgLen + 4 * ((C + M.tStgCpb - 1) / M.tStgCpb + 1) -
                  4 * sum_eoc;
    b.tracer += d * L * T * (C + 2 * C + E) + d * L * C;           // tlo rfac anti, yh
    b.tracer += d * L * T * C * touches + (S.stage > 0 && S.stage < 3 ? d * L * C : 0);
    if (S.use_hmix) b.tracer += 2 * d * E + d * L * E + d * L * T * C;  // dv dc h_edge xtr
  }
  b.vertex = 44 * V + 2 * d * V + d * E + d * L * (E + C + 2 * V);
  b.edge_diag = 16 * E + d * L * (C + V + 2 * E);
  if (S.ocean) {  // tracer-only run: no momentum, thickness or diagnostics
    b.edges = b.vertex = b.edge_diag = 0;
    b.cells -= d * E + d * L * E + d * L * C * (touches + 1 + (S.use_visc ? 1 : 0));
  }
  return b;
}

// One RK stage, in order on the context's stream: momentum, high-order
// flux (FCT), cells, tracer update (FCT), then the diagnostics of the
// stage's output.  (Running the independent pairs on two streams at once was
// measured slower on the level-8 workload: the pairs contend for L2 and HBM.)
// ctx (may be null) only times the launches while profiling is on.
template <int NT, int W, int WE, int LOL>
int launch_stage_t(mpas_ctx *ctx, hipStream_t st, const DMesh &M, const StageIO &S0,
                   const Bytes &b, double *vort_y, double *pv_vertex, double *h_edge_y,
                   double *pv_edge_y) {
  if (!LOL && M.nL != 1)  // the one-level kernels take level 0 only
    MPAS_FAIL(MPAS_EINVAL, "one-level stage kernels launched for a multi-level context");
  const StageIO S = LOL ? with_layout(S0, M) : S0;
  const int64_t lv = LOL ? M.nL : 1;  // lanes per entity and tracer lane group
  // level pairs: half the lanes, two levels each
  const bool pairs = LOL && MPAS_LEVEL_PAIRS && M.nL % 2 == 0;
  StageIO S2 = S;
  S2.lpl = 2;
  S2.nchunk = M.nL / 2;
  const bool fct = NT > 0 && S.scheme == MPAS_TRACER_FCT;
#define MPAS_STEP(expr) \
  do {                  \
    if (int rc = (expr)) return rc; \
  } while (0)
  bool merged_eh = false;  // momentum and high-order flux in one launch
  if constexpr (LOL == 0 && WE <= 12 && NT > 0) {
    if (MPAS_MERGE_EH && !S.ocean && fct && M.e2eLoc && M.a2cLoc && BS_EDGES == BS_HO) {
      MPAS_STEP(prof_begin(ctx, "k_stage_edges_ho", b.edges + b.ho));
      const int eg = xcd_grid(M.nE, BS_EDGES);
      const int hg = xcd_grid(M.nE * lanes_per_cell<NT, MPAS_TPL_HO>(), BS_HO);
      const size_t le = size_t(3) * M.eStgMax * sizeof(double);
      const size_t lh = size_t(M.aStgMax) * NT * sizeof(double);
      hipLaunchKernelGGL((k_stage_edges_ho<WE, NT>), dim3(eg + hg), dim3(BS_EDGES),
                         le > lh ? le : lh, st, M, S, eg);
      MPAS_STEP(after_launch("k_stage_edges_ho"));
      MPAS_STEP(prof_end(ctx));
      merged_eh = true;
    }
  }
  if (!S.ocean && !merged_eh) {
    MPAS_STEP(prof_begin(ctx, "k_stage_edges", b.edges));
    if (pairs) {
      hipLaunchKernelGGL((k_stage_edges<WE, LOL, 2>), dim3(xcd_grid(M.nE * (lv / 2), BS_EDGES)),
                         dim3(BS_EDGES), 0, st, M, S2);
    } else {
      bool staged = false;
      if constexpr (LOL == 0 && WE <= 12) {
        if (M.e2eLoc) {
          hipLaunchKernelGGL((k_stage_edges_lds<WE>), dim3(xcd_grid(M.nE, BS_EDGES)),
                             dim3(BS_EDGES), size_t(3) * M.eStgMax * sizeof(double), st, M, S);
          staged = true;
        }
      }
      if (!staged)
        hipLaunchKernelGGL((k_stage_edges<WE, LOL>), dim3(xcd_grid(M.nE * lv, BS_EDGES)),
                           dim3(BS_EDGES), 0, st, M, S);
    }
    MPAS_STEP(after_launch("k_stage_edges"));
    MPAS_STEP(prof_end(ctx));
  }
  if constexpr (NT > 0) {
    if (fct && !merged_eh) {
      MPAS_STEP(prof_begin(ctx, "k_stage_ho", b.ho));
      if (pairs)
        hipLaunchKernelGGL((k_stage_ho<NT, WE, LOL, 2>),
                           dim3(xcd_grid(M.nE * (lv / 2) * lanes_per_cell<NT, MPAS_TPL_HO>(),
                                         BS_HO)),
                           dim3(BS_HO), 0, st, M, S2);
      else {
        bool staged = false;
        if constexpr (LOL == 0 && WE <= 12) {
          if (M.a2cLoc) {
            hipLaunchKernelGGL((k_stage_ho_lds<NT, WE>),
                               dim3(xcd_grid(M.nE * lanes_per_cell<NT, MPAS_TPL_HO>(), BS_HO)),
                               dim3(BS_HO), size_t(M.aStgMax) * NT * sizeof(double), st, M, S);
            staged = true;
          }
        }
        if (!staged)
          hipLaunchKernelGGL((k_stage_ho<NT, WE, LOL>),
                             dim3(xcd_grid(M.nE * lv * lanes_per_cell<NT, MPAS_TPL_HO>(), BS_HO)),
                             dim3(BS_HO), 0, st, M, S);
      }
      MPAS_STEP(after_launch("k_stage_ho"));
      MPAS_STEP(prof_end(ctx));
    }
  }
  MPAS_STEP(prof_begin(ctx, "k_stage_cells", b.cells));
  {
    const dim3 grid(xcd_grid(M.nC * lv * lanes_per_cell<NT, MPAS_TPL_CELLS>(), BS_CELLS));
    if (NT == 0 || S.scheme == MPAS_TRACER_NONE)
      hipLaunchKernelGGL((k_stage_cells<NT, W, LOL, MPAS_TRACER_NONE>), grid, dim3(BS_CELLS), 0,
                         st, M, S);
    else if (S.scheme == MPAS_TRACER_CENTRED)
      hipLaunchKernelGGL((k_stage_cells<NT, W, LOL, MPAS_TRACER_CENTRED>), grid, dim3(BS_CELLS),
                         0, st, M, S);
    else {
      bool staged = false;
      if constexpr (LOL == 0 && NT > 0) {
        if (M.c2eLoc) {
          hipLaunchKernelGGL((k_stage_cells_lds<NT, W>), grid, dim3(BS_CELLS),
                             size_t(M.stgMax) * (5 + NT) * sizeof(double), st, M, S);
          staged = true;
        }
      }
      if constexpr (LOL == 0 && NT > 0 && NT % 2 == 0 && W <= 6 &&
                    tracers_per_lane<NT>() % 2 == 0 && MPAS_TPL_CELLS == 0) {
        if (!staged && M.c2cLocC) {
          hipLaunchKernelGGL((k_stage_cells_cst<NT, W>), grid, dim3(BS_CELLS),
                             size_t(M.cStgMax) * (NT + 1) * sizeof(double), st, M, S);
          staged = true;
        }
      }
      if (!staged)
        hipLaunchKernelGGL((k_stage_cells<NT, W, LOL, MPAS_TRACER_FCT>), grid, dim3(BS_CELLS), 0,
                           st, M, S);
    }
  }
  MPAS_STEP(after_launch("k_stage_cells"));
  MPAS_STEP(prof_end(ctx));
  bool merged = false;  // the vertex diagnostics went with the tracer update
  if constexpr (NT > 0) {
    if (fct) {
      const dim3 grid(xcd_grid(M.nC * lv * lanes_per_cell<NT, tpl_tracer(LOL)>(), BS_TRACER));
      const size_t occ =
          occupancy_lds(LOL ? MPAS_WG_PER_CU_TRACER_LANES : MPAS_WG_PER_CU_TRACER);
      bool staged = false;
      if constexpr (LOL == 0 && W <= 6 && NT % 2 == 0) {
        if (M.c2cLoc) {
          const size_t bytes = size_t(3) * NT * M.tStgMax * sizeof(double);
          const size_t occ_s = occupancy_lds(MPAS_WG_PER_CU_TRACER_LDS);
          const size_t lds = bytes > occ_s ? bytes : occ_s;
          if (MPAS_MERGE_TV && !S.ocean) {
            MPAS_STEP(prof_begin(ctx, "k_stage_tracer_vertex", b.tracer + b.vertex));
            const int vgrid = xcd_grid(M.nV, BS_DIAG);
            hipLaunchKernelGGL((k_stage_tracer_vertex<NT, W>), dim3(grid.x + vgrid),
                               dim3(BS_TRACER), lds, st, M, S, int(grid.x), vort_y, pv_vertex);
            merged = true;
          } else {
            MPAS_STEP(prof_begin(ctx, "k_stage_tracer", b.tracer));
            hipLaunchKernelGGL((k_stage_tracer_lds<NT, W>), grid, dim3(BS_TRACER), lds, st, M, S);
          }
          staged = true;
        }
      }
      if (!staged) {
        MPAS_STEP(prof_begin(ctx, "k_stage_tracer", b.tracer));
        hipLaunchKernelGGL((k_stage_tracer<NT, W, LOL>), grid, dim3(BS_TRACER), occ, st, M, S);
      }
      MPAS_STEP(after_launch("k_stage_tracer"));
      MPAS_STEP(prof_end(ctx));
    }
  }
  if (S.ocean) return 0;  // u and h do not change: no diagnostics
  if (!merged) MPAS_STEP(prof_begin(ctx, "k_diag_vertex", b.vertex));
  if (merged) {
  } else if (pairs)
    hipLaunchKernelGGL((k_diag_vertex<LOL, 2>), dim3(xcd_grid(M.nV * (lv / 2), BS_DIAG)),
                       dim3(BS_DIAG), 0, st, M, S2, S.yh, S.yu, vort_y, pv_vertex);
  else
    hipLaunchKernelGGL((k_diag_vertex<LOL>), dim3(xcd_grid(M.nV * lv, BS_DIAG)), dim3(BS_DIAG),
                       0, st, M, S, S.yh, S.yu, vort_y, pv_vertex);
  if (!merged) {
    MPAS_STEP(after_launch("k_diag_vertex"));
    MPAS_STEP(prof_end(ctx));
  }
  MPAS_STEP(prof_begin(ctx, "k_diag_edge", b.edge_diag));
  if (pairs)
    hipLaunchKernelGGL((k_diag_edge<LOL, 2>), dim3(xcd_grid(M.nE * (lv / 2), BS_DIAG)),
                       dim3(BS_DIAG), 0, st, M, S2, S.yh, pv_vertex, h_edge_y, pv_edge_y);
  else
    hipLaunchKernelGGL((k_diag_edge<LOL>), dim3(xcd_grid(M.nE * lv, BS_DIAG)), dim3(BS_DIAG), 0,
                       st, M, S, S.yh, pv_vertex, h_edge_y, pv_edge_y);
  MPAS_STEP(after_launch("k_diag_edge"));
  MPAS_STEP(prof_end(ctx));
#undef MPAS_STEP
  return 0;
}

#define MPAS_STAGE_ARGS ctx, st, M, S, b, vort_y, pv_vertex, h_edge_y, pv_edge_y
template <int NT>
int launch_stage_w(mpas_ctx *ctx, hipStream_t st, const DMesh &M, const StageIO &S,
                   const Bytes &b, double *vort_y, double *pv_vertex, double *h_edge_y,
                   double *pv_edge_y) {
  if (layout_of(M)) {
    if (M.narrow) return launch_stage_t<NT, 6, 10, 1>(MPAS_STAGE_ARGS);
    return launch_stage_t<NT, 8, 16, 1>(MPAS_STAGE_ARGS);
  }
  if (M.narrow) return launch_stage_t<NT, 6, 10, 0>(MPAS_STAGE_ARGS);
  return launch_stage_t<NT, 8, 16, 0>(MPAS_STAGE_ARGS);
}

int launch_stage_nt(mpas_ctx *ctx, hipStream_t st, const DMesh &M, const StageIO &S,
                    const Bytes &b, double *vort_y, double *pv_vertex, double *h_edge_y,
                    double *pv_edge_y) {
  switch (M.nT) {
    case 0: return launch_stage_w<0>(MPAS_STAGE_ARGS);
    case 1: return launch_stage_w<1>(MPAS_STAGE_ARGS);
    case 2: return launch_stage_w<2>(MPAS_STAGE_ARGS);
    case 3: return launch_stage_w<3>(MPAS_STAGE_ARGS);
    case 4: return launch_stage_w<4>(MPAS_STAGE_ARGS);
    case 5: return launch_stage_w<5>(MPAS_STAGE_ARGS);
    case 6: return launch_stage_w<6>(MPAS_STAGE_ARGS);
    case 7: return launch_stage_w<7>(MPAS_STAGE_ARGS);
    case 8: return launch_stage_w<8>(MPAS_STAGE_ARGS);
  }
  MPAS_FAIL(MPAS_EINVAL, "the fused step supports at most 8 tracers");
}

}  // namespace

StageIO stage_params(const mpas_ctx *ctx, int s, double dt) {
  const DMesh &M = ctx->dm;
  const mpas_params &P = ctx->params;
  StageIO S{};
  S.stage = s;
  S.scheme = M.nT ? P.tracer_scheme : MPAS_TRACER_NONE;
  S.freeze_u = P.freeze_u;
  S.use_visc = P.visc2 != 0.0;
  S.use_hmix = M.nT && P.kappa2 != 0.0;
  S.w_rk = rk_weight(s, dt);
  S.w_sub = rk_substep_weight(s, dt);
  S.dt = dt;
  S.visc_coef = M.msd * P.visc2;
  S.kappa = P.kappa2;
  return S;
}

int launch_stage(mpas_ctx *ctx, const StageIO &S, double *vort_y, double *pv_vertex,
                 double *h_edge_y, double *pv_edge_y) {
  const DMesh &M = ctx->dm;
  const Bytes b = stage_bytes(M, S, ctx->sum_eoe, ctx->sum_adv, ctx->sum_eoc);
  return launch_stage_nt(ctx, ctx->stream, M, S, b, vort_y, pv_vertex, h_edge_y, pv_edge_y);
}

int launch_stage_dm(const DMesh &M, hipStream_t st, const StageIO &S, double *vort_y,
                    double *pv_vertex, double *h_edge_y, double *pv_edge_y) {
  return launch_stage_nt(nullptr, st, M, S, Bytes{}, vort_y, pv_vertex, h_edge_y, pv_edge_y);
}

int diagnostics_dm(const DMesh &M, hipStream_t st, int use_visc, const double *u,
                   const double *h, double *ke, double *div, double *vort, double *pv_vertex,
                   double *h_edge, double *pv_edge) {
  hipLaunchKernelGGL(k_diag_cell, dim3(xcd_grid(M.nC, BS)), dim3(BS), 0, st, M, u, ke, div,
                     use_visc);
  if (int rc = after_launch("k_diag_cell")) return rc;
  const StageIO S = with_layout(StageIO{}, M);
  if (layout_of(M)) {
    hipLaunchKernelGGL(k_diag_vertex<1>, dim3(xcd_grid(int64_t(M.nV) * M.nL, BS_DIAG)),
                       dim3(BS_DIAG), 0, st, M, S, h, u, vort, pv_vertex);
  } else {
    hipLaunchKernelGGL(k_diag_vertex<0>, dim3(xcd_grid(M.nV, BS_DIAG)), dim3(BS_DIAG), 0, st, M,
                       S, h, u, vort, pv_vertex);
  }
  if (int rc = after_launch("k_diag_vertex")) return rc;
  if (layout_of(M)) {
    hipLaunchKernelGGL(k_diag_edge<1>, dim3(xcd_grid(int64_t(M.nE) * M.nL, BS_DIAG)),
                       dim3(BS_DIAG), 0, st, M, S, h, pv_vertex, h_edge, pv_edge);
  } else {
    hipLaunchKernelGGL(k_diag_edge<0>, dim3(xcd_grid(M.nE, BS_DIAG)), dim3(BS_DIAG), 0, st, M, S,
                       h, pv_vertex, h_edge, pv_edge);
  }
  return after_launch("k_diag_edge");
}

int diagnostics_into(mpas_ctx *ctx, const double *u, const double *h, double *ke, double *div,
                     double *vort, double *pv_vertex, double *h_edge, double *pv_edge) {
  return diagnostics_dm(ctx->dm, ctx->stream, ctx->params.visc2 != 0.0, u, h, ke, div, vort,
                        pv_vertex, h_edge, pv_edge);
}

// Diagnostics of the state in (u, h) into ping-pong slot `slot`
// (compute_solve_diagnostics, plus the divergence the viscosity reads).
int sw_diagnostics(mpas_ctx *ctx, const double *u, const double *h, int slot) {
  SwState &S = *ctx->sw;
  return diagnostics_into(ctx, u, h, S.ke[slot], S.div[slot], S.vort[slot], S.pv_vertex,
                          S.h_edge[slot], S.pv_edge[slot]);
}

// nsteps of rk4_timestep (apps/ocean/run.cc:10-50): the FCT tendency of
// the tracers under the fixed u (normalVelocity), h (layerThickness) and
// h_edge (layerThicknessEdge, init_functions.cc:30-45, the h_edge the
// state upload formed), ping-pong buffers for the stage tracers (prev).
int ocean_rk4_steps(mpas_ctx *ctx, double dt, int nsteps) {
  SwState &W = *ctx->sw;
  for (int step = 0; step < nsteps; step++) {
    for (int s = 0; s < 4; s++) {
      const int cur = s % 2, nxt = 1 - cur;
      StageIO S = stage_params(ctx, s, dt);
      S.ocean = 1;
      S.xu = W.u;
      S.xh = W.h;
      S.xtr = s == 0 ? W.tr : W.ptr[cur];
      S.h_edge = W.h_edge[W.cur];
      S.h_s = W.h_s;
      S.tr0 = W.tr;
      S.ntr = W.ntr;
      S.ytr = s < 3 ? W.ptr[nxt] : W.tr;
      S.anti = W.anti;
      S.rfac = W.rfac;
      S.tlo = W.tlo;
      if (int rc = launch_stage(ctx, S, nullptr, nullptr, nullptr, nullptr)) return rc;
    }
    W.elapsed += dt;
  }
  return 0;
}

// The buffers stage s reads and writes (apps/sw/run.cc:13-70): the
// reference's copy tasks become these roles, with the ping-pong slot W.cur
// holding the stage's input diagnostics.
StageIO sw_stage_io(const mpas_ctx *ctx, const SwState &W, int s, double dt) {
  const int cur = W.cur, nxt = 1 - W.cur;
  StageIO S = stage_params(ctx, s, dt);
  S.xu = s == 0 ? W.u : W.pu[cur];
  S.xh = s == 0 ? W.h : W.ph[cur];
  S.xtr = s == 0 ? W.tr : W.ptr[cur];
  S.h_edge = W.h_edge[cur];
  S.ke = W.ke[cur];
  S.pv_edge = W.pv_edge[cur];
  S.vort = W.vort[cur];
  S.div = W.div[cur];
  S.h_s = W.h_s;
  S.u0 = W.u;
  S.h0 = W.h;
  S.tr0 = W.tr;
  S.nu = W.nu;
  S.nh = W.nh;
  S.ntr = W.ntr;
  S.yu = s < 3 ? W.pu[nxt] : W.u;
  S.yh = s < 3 ? W.ph[nxt] : W.h;
  S.ytr = s < 3 ? W.ptr[nxt] : W.tr;
  S.ke_y = W.ke[nxt];
  S.div_y = W.div[nxt];
  S.anti = W.anti;
  S.rfac = W.rfac;
  S.tlo = W.tlo;
  return S;
}

// nsteps of apps/sw/run.cc:13-70.
int sw_rk4_steps(mpas_ctx *ctx, double dt, int nsteps, double *const *frozen_rfac) {
  SwState &W = *ctx->sw;
  for (int step = 0; step < nsteps; step++) {
    for (int s = 0; s < 4; s++) {
      const int nxt = 1 - W.cur;
      StageIO S = sw_stage_io(ctx, W, s, dt);
      if (frozen_rfac) {
        S.rfac = frozen_rfac[4 * step + s];
        S.keep_rfac = 1;
      }
      if (int rc = launch_stage(ctx, S, W.vort[nxt], W.pv_vertex, W.h_edge[nxt], W.pv_edge[nxt]))
        return rc;
      W.cur = nxt;
    }
    W.elapsed += dt;
  }
  return 0;
}

// The same on buffers the caller describes, without the context's
// profiling (which creates events): no allocation of any kind.
int sw_rk4_steps_on(const mpas_ctx *ctx, SwState &W, double dt, int nsteps) {
  for (int step = 0; step < nsteps; step++) {
    for (int s = 0; s < 4; s++) {
      const int nxt = 1 - W.cur;
      const StageIO S = sw_stage_io(ctx, W, s, dt);
      if (int rc = launch_stage_dm(ctx->dm, ctx->stream, S, W.vort[nxt], W.pv_vertex,
                                   W.h_edge[nxt], W.pv_edge[nxt]))
        return rc;
      W.cur = nxt;
    }
    W.elapsed += dt;
  }
  return 0;
}

}  // namespace mpasx
