// Tangent-linear model (TLM) and discrete adjoint of the stage-fused RK4
// step with FCT tracers (the forward is apps/sw/run.cc:13-70 with the
// tracer tasks of apps/ocean/tasks/solve.cc:40-366 as the tendency; the
// reference has no derivative code).
//
// Linearisation.  One stage maps X = (u, h, tracers) to the tendencies
// (tend_u, tend_h, h-weighted tracer tendency); the RK4 bookkeeping combines
// them (u and h linearly, tracers as (h0*T0 + w*tend)/h).  The FCT limiter
// factors R+/R- of every cell are frozen at the forward pass's values; every
// branch (upwind side max/min(F, 0), the sign of F in the 3rd-order
// coefficient, the sign of the antidiffusive flux that selects which factor
// pair scales it) follows the forward state.  With these fixed the stage is
// a smooth map, and TLM and adjoint are its derivative and the transpose of
// that derivative.
//
// Gathers only.  Each forward scatter (a cell or vertex value fed by its
// edges) becomes, in the adjoint, a gather over the transposed list: the
// edges that hold an edge in their edgesOnEdge (with that slot's weight),
// the edges whose advection stencil holds a cell (with the stencil
// coefficients), and the vertices of a cell (with the kite area the vertex
// pairs with that cell).  No atomics; every value is summed in one fixed
// order, so both models are deterministic.
//
// Tape.  The forward keeps a checkpoint of the state every K steps.  The
// backward sweep replays each segment with the same fused stage kernels,
// writing every stage's input state, diagnostics, limiter factors and
// antidiffusive fluxes into a tape of 4K + 1 stage records (the replay is
// bit-identical to the first pass), then sweeps the segment in reverse.
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "ctx.h"
#include "stage.h"

namespace mpasx {

// Forward record of one RK stage: its input state X, the diagnostics of X,
// and the limiter factors and antidiffusive fluxes the stage formed.
struct Rec {
  double *u, *h, *tr, *he, *ke, *pve, *vort, *div, *rfac, *anti;
};

// Transposed lists, slot-major ELL with run-time widths.
struct AdjMesh {
  // edges e2 whose edgesOnEdge holds e, and weightsOnEdge(e2, that slot)
  const uint8_t *nEt;
  const int *et;
  const double *wt;
  // where edgesOnEdge is symmetric: nibble j, the slot of e in the row of
  // edgesOnEdge(e, j) (0 past the row)
  const uint64_t *rslot;
  // (edge, slot) pairs whose advCellsForEdge entry is c, as
  // slot << kAtSlotShift | edge
  const uint8_t *nAt;
  const int *at;
  // vertices of c with the kite area that h_vertex pairs with c (the i-th
  // ascending cell of the vertex takes the i-th file-order kite area)
  const uint8_t *nVc;
  const int *vc;
  const double *vk;
  int wVc, wAt;  // padded widths of the vc and at rows
  // elements of one stencil slot's part of hbf (edges x levels x tracers):
  // each slot's part is addressed from its own base, so that every 32-bit
  // offset stays within one part while the whole may pass 4 GiB
  uint64_t hslot;
};
constexpr int kAtSlotShift = 24;  // edges per mesh below 2^24

struct AdjState {
  int K = 0;           // steps per checkpoint segment
  int nsteps = 0;      // steps of the last mpas_adj_forward
  double dt = 0;
  bool forward_done = false;
  int taped = 0;       // steps from step 0 whose records the tape holds
  int tape_seg = -1;   // checkpoint segment whose records the tape holds
  std::vector<Rec> rec;          // 4K + 1 stage records
  std::vector<double *> ck;      // checkpoints: (u, h, tr) per segment
  double *pv_vertex = nullptr, *nu = nullptr, *nh = nullptr, *ntr = nullptr, *tlo = nullptr;
  // tangent-linear perturbations of the 5 stage states of a step, the
  // RK4 accumulators, and per-stage intermediates
  double *du[5] = {}, *dh[5] = {}, *dT[5] = {};
  double *dnu = nullptr, *dnh = nullptr, *dQ = nullptr;
  double *dvort = nullptr, *dpvv = nullptr, *dhe = nullptr, *dpve = nullptr;
  double *dke = nullptr, *ddiv = nullptr, *dflux = nullptr;
  // adjoint: state adjoints (ping-pong), the adjoints of the step's final
  // state after the final division (aw_*), the current stage's tendency
  // adjoints (tb_*), and per-stage intermediates
  double *xb_u[2] = {}, *xb_h[2] = {}, *xb_tr[2] = {};
  double *aw_u = nullptr, *aw_h = nullptr, *aw_tr = nullptr;
  double *tb_u = nullptr, *tb_h = nullptr, *tb_tr = nullptr;
  double *ubA = nullptr, *heb = nullptr, *pveb = nullptr, *r1 = nullptr, *r2 = nullptr,
         *hbf = nullptr, *vT = nullptr, *hvT = nullptr, *keT = nullptr, *divT = nullptr;
  AdjMesh T{};
  std::vector<void *> bufs;
  // A second tape with its own replay buffers: the backward sweep replays
  // segment m - 1 into it on a second stream while it sweeps segment m from
  // the first (the replay needs only the checkpoint; the sweep reads only the
  // tape and the adjoint buffers), then the two swap.  Allocated at the
  // first sweep of more than one segment when the device has the memory;
  // without it the replays run in turn on the context's stream.
  struct Tape {
    std::vector<Rec> rec;
    double *pv_vertex = nullptr, *nu = nullptr, *nh = nullptr, *ntr = nullptr,
           *tlo = nullptr;
  } spare;
  int spare_state = 0;  // 0 not tried, 1 allocated, -1 not enough memory
  hipStream_t replay_stream = nullptr;
  hipEvent_t ev_swept = nullptr, ev_replayed = nullptr;
  // tile and schedule: workgroup sizes of the adjoint (k_ad_*) and
  // tangent-linear (k_tl_*) kernels; replays beside the sweep (spare tape)
  // or in turn; each segment's reverse sweep replayed from a hipGraph
  // (captured per tape, parity and length) or launched directly
  int ad_block = 256, tl_block = 256, overlap = 1, graph = 0;
  // the stage-0 kernels of each step but a segment's last in the sweep run
  // the next step's prep (k_ad_prep otherwise: MPAS_HIP_ADJ_FUSE_PREP=0
  // when the context is made); the same bits either way
  int fuse_prep = 1;
  // 0: sweeps are never replayed from graphs (MPAS_HIP_ADJ_GRAPH=0 when the
  // context is made; mpas_adj_config's graph is then ignored and the
  // autotuner does not try it), for tools that cannot follow graph capture
  // (rocprofv3's kernel tracing stops the process in it)
  int graph_ok = 1;
  // the edge part reads the slot-paired stencils (k_ad_edge_a's FAST form;
  // MPAS_HIP_ADJ_PAIRED=0 when the context is made keeps the slot-major rows)
  int edge_fast = 0;
  // with fewer than kOneLaneLevels levels the vertex part runs in the cell
  // part's launch, whose cells form their vertices' h_vertex adjoints
  // themselves (MPAS_HIP_ADJ_MERGE_VERTEX=0 when the context is made keeps
  // two launches; the same bits either way)
  int merge_vertex = 1;
  // with a multiple of 64 levels the edge and cell parts take their entity
  // from the wave's first lane and load its static rows as scalars
  // (MPAS_HIP_ADJ_UNIFORM=0 when the context is made keeps vector loads;
  // the same bits either way)
  int uni_ok = 1;
  std::map<std::tuple<const void *, const void *, int, int>, hipGraphExec_t> sweeps;
};

void drop_sweeps(AdjState &A) {
  for (auto &kv : A.sweeps) (void)hipGraphExecDestroy(kv.second);
  A.sweeps.clear();
}

void adj_free(mpas_ctx *ctx) {
  if (!ctx->adj) return;
  AdjState *A = ctx->adj;
  if (A->replay_stream) {
    (void)hipStreamSynchronize(A->replay_stream);
    (void)hipStreamDestroy(A->replay_stream);
  }
  drop_sweeps(*A);
  if (A->ev_swept) (void)hipEventDestroy(A->ev_swept);
  if (A->ev_replayed) (void)hipEventDestroy(A->ev_replayed);
  for (void *p : A->bufs) (void)hipFree(p);
  delete A;
  ctx->adj = nullptr;
}

namespace {

// Workgroup sizes are chosen at run time (64 to 512; mpas_adj_config,
// mpas_adj_autotune): the kernels are built for the largest.
constexpr int kAdjMaxBlock = 512;

// Per-stage constants of the linearised kernels.
struct AdjParams {
  int s, nT, freeze_u, use_visc, use_hmix, fct;
  double w_rk, w_sub;      // weights of stage s
  double w_rk_m, w_sub_m;  // weights of stage s - 1 (adjoint bookkeeping)
  double visc_coef, kappa;
};

__device__ __forceinline__ double fsign(double F) { return (F < 0.0) ? -1.0 : double(F > 0.); }

// An advection transpose entry (slot << kAtSlotShift | edge): its edge, and
// the base of its slot's part of hbf.
__device__ __forceinline__ uint32_t at_edge(uint32_t a) {
  return a & ((1u << kAtSlotShift) - 1u);
}
__device__ __forceinline__ const double *hbf_at(const AdjMesh &A, const double *hbf, uint32_t a) {
  return hbf + uint64_t(a >> kAtSlotShift) * A.hslot;
}

// Loads of the static mesh and transpose arrays.  U: the wave holds one
// entity (levels on lanes, a multiple of 64 levels) whose index was taken
// from the wave's first lane, so the index is wave-uniform and the
// constant-space load (ldc) becomes a scalar load, off the vector memory
// path; otherwise an ordinary global load.
template <bool U, class T>
__device__ __forceinline__ T ldm(const T *__restrict__ p, uint32_t i) {
  if constexpr (U) return ldc(p, i);
  else return ld(p, i);
}
template <bool U, class V, class S>
__device__ __forceinline__ V ldmv(const S *__restrict__ p, uint32_t i) {
  if constexpr (U) return ldc_vec<V>(p, i);
  else return ld_vec<false, V>(p, i);
}

// Work item -> (entity, level): one level per lane, an entity's levels on
// adjacent lanes (with one level, the entity is the item).  groups > 1
// keeps `groups` adjacent lanes per (entity, level), lane = their index.
template <int groups = 1>
__device__ __forceinline__ int level_item_at(int item, uint32_t L, uint32_t &k0, uint32_t &k1,
                                             int &lane) {
  const int q = item / groups;
  lane = item - q * groups;
  const int ent = q / int(L);
  k0 = uint32_t(q - ent * int(L));
  k1 = k0 + 1;
  return ent;
}
template <int groups = 1>
__device__ __forceinline__ int level_item(uint32_t L, uint32_t &k0, uint32_t &k1, int &lane) {
  return level_item_at<groups>(xcd_item(), L, k0, k1, lane);
}

// Slope of the rescaled antidiffusive flux in the flux x with the limiter
// factors frozen: max(0,x)*min(out1,in2) + min(0,x)*min(in1,out2)
// (ocean/solve.cc:270-282).
__device__ __forceinline__ double slope_of(double in1, double out1, double in2, double out2,
                                           double x) {
  return x > 0.0 ? smin(out1, in2) : (x < 0.0 ? smin(in1, out2) : 0.0);
}
__device__ __forceinline__ double frozen_slope(const double *__restrict__ rfac, uint32_t i1,
                                               uint32_t i2, double x) {
  const double in1 = ld(rfac, 2 * i1), out1 = ld(rfac, 2 * i1 + 1);
  const double in2 = ld(rfac, 2 * i2), out2 = ld(rfac, 2 * i2 + 1);
  return slope_of(in1, out1, in2, out2, x);
}

// h_vertex and pv_vertex of the recorded state, formed as the forward forms
// them (ascending cells with file-order kite areas).
__device__ __forceinline__ void vertex_state(const DMesh &M, int v, uint32_t k, const Rec &R,
                                             const uint32_t (&c)[3], const double (&kt)[3],
                                             double at, double &hv, double &pvv) {
  const uint32_t L = M.nL;
  double s = 0.0;
  s += ld(R.h, c[0] * L + k) * kt[0];
  s += ld(R.h, c[1] * L + k) * kt[1];
  s += ld(R.h, c[2] * L + k) * kt[2];
  hv = s / at;
  pvv = (ld(M.fV, v) + ld(R.vort, v * L + k)) / hv;
}

// ---- tangent-linear model -----------------------------------------------------

__global__ __launch_bounds__(kAdjMaxBlock) void k_tl_vertex(DMesh M, AdjParams P, Rec R,
                                                  const double *__restrict__ du,
                                                  const double *__restrict__ dh, double *dvort,
                                                  double *dpvv) {
  uint32_t k0, k1;
  int lane_;
  const int v = level_item(M.nL, k0, k1, lane_);
  if (v >= M.nV) return;
  const uint32_t L = M.nL, V = M.nV;
  int x[3];
  uint32_t c[3];
  double kt[3];
  for (int i = 0; i < 3; i++) {
    x[i] = ld(M.v2e, i * V + v);
    c[i] = ld(M.v2c, i * V + v);
    kt[i] = ld(M.kite, i * V + v);
  }
  const double at = ld(M.areaTri, v);
  for (uint32_t k = k0; k < k1; k++) {
    double dcirc = 0.0, dhv = 0.0;
    for (int i = 0; i < 3; i++) {
      const uint32_t e = v2e_id(x[i]);
      dcirc += v2e_sign(x[i]) * (ld(M.dc, e) * ld(du, e * L + k));
      dhv += ld(dh, c[i] * L + k) * kt[i];
    }
    const double dvo = dcirc / at;
    dhv /= at;
    double hv, pvv;
    vertex_state(M, v, k, R, c, kt, at, hv, pvv);
    st(dvort, v * L + k, dvo);
    st(dpvv, v * L + k, (dvo - pvv * dhv) / hv);
  }
}

// h_edge and pv_edge perturbations, and the perturbation of the edge's net
// tracer flux into its cells: (1 - lambda) dLO + lambda dHO - dH (upwind,
// 3rd-order, del2 mixing flux).
__global__ __launch_bounds__(kAdjMaxBlock) void k_tl_edge_a(DMesh M, AdjParams P, Rec R,
                                                  const double *__restrict__ du,
                                                  const double *__restrict__ dh,
                                                  const double *__restrict__ dT,
                                                  const double *__restrict__ dpvv, double *dhe,
                                                  double *dpve, double *dflux) {
  uint32_t k0, k1;
  int lane_;
  const int e = level_item(M.nL, k0, k1, lane_);
  if (e >= M.nE) return;
  const uint32_t L = M.nL, E = M.nE, NT = P.nT;
  const int2 cc = ld(M.e2c, e), vv = ld(M.e2v, e);
  const uint32_t c1 = cc.x, c2 = cc.y;
  const double dv = ld(M.dv, e), dc = ld(M.dc, e);
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t o = e * L + k;
    const double dhe_ = 0.5 * (ld(dh, c1 * L + k) + ld(dh, c2 * L + k));
    st(dhe, o, dhe_);
    st(dpve, o, 0.5 * ld(dpvv, uint32_t(vv.x) * L + k) + 0.5 * ld(dpvv, uint32_t(vv.y) * L + k));
    if (!P.fct) continue;
    const double ue = ld(R.u, o), hee = ld(R.he, o);
    const double F = ue * hee, dF = ld(du, o) * hee + ue * dhe_;
    const double sgn = fsign(F);
    const double Fp = F > 0.0 ? F : 0.0, Fm = F < 0.0 ? F : 0.0;
    const double dFp = F > 0.0 ? dF : 0.0, dFm = F < 0.0 ? dF : 0.0;
    const int na = ld(M.nAdv, e);
    const double rho = M.msd * P.kappa * dv / dc;
    for (uint32_t t = 0; t < NT; t++) {
      const uint32_t i1 = (c1 * L + k) * NT + t, i2 = (c2 * L + k) * NT + t;
      const double T1 = ld(R.tr, i1), T2 = ld(R.tr, i2);
      const double dT1 = ld(dT, i1), dT2 = ld(dT, i2);
      const double dlo = dv * (dFp * T1 + Fp * dT1 + dFm * T2 + Fm * dT2);
      double dho = 0.0;
      for (int i = 0; i < na; i++) {
        const uint32_t ci = ld(M.adv, i * E + e);
        const double cf = ld(M.advCoef, i * E + e) + ld(M.advCoef3, i * E + e) * M.coef3rd * sgn;
        const uint32_t ii = (ci * L + k) * NT + t;
        dho += (ld(dT, ii) * F + ld(R.tr, ii) * dF) * cf;
      }
      const double lam = frozen_slope(R.rfac, i1, i2, ld(R.anti, o * NT + t));
      double dphi = (1.0 - lam) * dlo + lam * dho;
      if (P.use_hmix) dphi -= rho * (dhe_ * (T2 - T1) + hee * (dT2 - dT1));
      st(dflux, o * NT + t, dphi);
    }
  }
}

// Cell part: ke / divergence perturbations, tend_h and the tracer tendency,
// then the RK4 bookkeeping of h and the tracers (solve.cc:297-305,
// 332-340, 350-357 differentiated).
__global__ __launch_bounds__(kAdjMaxBlock) void k_tl_cell(DMesh M, AdjParams P, Rec R, Rec Ry, Rec R0,
                                                const double *__restrict__ du,
                                                const double *__restrict__ dh0,
                                                const double *__restrict__ dT0,
                                                const double *__restrict__ dhe,
                                                const double *__restrict__ dflux, double *dnh,
                                                double *dQ, double *dke, double *ddiv,
                                                double *dhy, double *dTy) {
  uint32_t k0, k1;
  int lane_;
  const int c = level_item(M.nL, k0, k1, lane_);
  if (c >= M.nC) return;
  const uint32_t L = M.nL, C = M.nC, NT = P.nT;
  const int n = ld(M.nEoC, c);
  const uint8_t neg = ld(M.sgnNeg, c);
  const double ar = ld(M.area, c);
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t ol = c * L + k;
    double dk = 0.0, dd = 0.0, dth = 0.0, dtT[8] = {};
    for (int j = 0; j < n; j++) {
      const uint32_t e = ld(M.c2e, j * C + c), oe = e * L + k;
      const double sg = sign_on_cell(neg, j), dv = ld(M.dv, e), dc = ld(M.dc, e);
      const double u = ld(R.u, oe), duj = ld(du, oe);
      dk += 0.5 * dc * dv * u * duj;
      dd += (-sg) * dv * duj;
      dth += sg * dv * (duj * ld(R.he, oe) + u * ld(dhe, oe));
      for (uint32_t t = 0; t < NT; t++) dtT[t] += sg * ld(dflux, oe * NT + t);
    }
    st(dke, ol, dk / ar);
    if (P.use_visc) st(ddiv, ol, dd / ar);
    dth /= ar;
    const double d0 = ld(dh0, ol);
    double dhy_;
    if (P.s == 0) {
      st(dnh, ol, d0 + P.w_rk * dth);
      dhy_ = d0 + P.w_sub * dth;
    } else if (P.s < 3) {
      st(dnh, ol, ld(dnh, ol) + P.w_rk * dth);
      dhy_ = d0 + P.w_sub * dth;
    } else {
      dhy_ = ld(dnh, ol) + P.w_rk * dth;
    }
    st(dhy, ol, dhy_);
    if (!NT) continue;
    const double hy = ld(Ry.h, ol), h0 = ld(R0.h, ol);
    for (uint32_t t = 0; t < NT; t++) {
      const uint32_t ot = ol * NT + t;
      const double te = dtT[t] / ar;
      const double Ty = ld(Ry.tr, ot), T0 = ld(R0.tr, ot), dt0 = ld(dT0, ot);
      double y;
      if (P.s == 0) {
        st(dQ, ot, dt0 * h0 + T0 * d0 + P.w_rk * te);
        y = (d0 * T0 + h0 * dt0 + P.w_sub * te - Ty * dhy_) / hy;
      } else if (P.s < 3) {
        st(dQ, ot, ld(dQ, ot) + P.w_rk * te);
        y = (d0 * T0 + h0 * dt0 + P.w_sub * te - Ty * dhy_) / hy;
      } else {
        y = (ld(dQ, ot) + P.w_rk * te - Ty * dhy_) / hy;
      }
      st(dTy, ot, y);
    }
  }
}

// tend_u perturbation (PV flux, ke and height gradients, del2 viscosity) and
// the RK4 bookkeeping of u.
__global__ __launch_bounds__(kAdjMaxBlock) void k_tl_edge_b(DMesh M, AdjParams P, Rec R,
                                                  const double *__restrict__ du,
                                                  const double *__restrict__ du0,
                                                  const double *__restrict__ dh,
                                                  const double *__restrict__ dhe,
                                                  const double *__restrict__ dpve,
                                                  const double *__restrict__ dke,
                                                  const double *__restrict__ ddiv,
                                                  const double *__restrict__ dvort, double *dnu,
                                                  double *duy) {
  uint32_t k0, k1;
  int lane_;
  const int e = level_item(M.nL, k0, k1, lane_);
  if (e >= M.nE) return;
  const uint32_t L = M.nL, E = M.nE;
  if (P.freeze_u) {  // u stays the step's initial u (solve.cc:307-313, 360-366)
    for (uint32_t k = k0; k < k1; k++) st(duy, e * L + k, ld(du0, e * L + k));
    return;
  }
  const int2 cc = ld(M.e2c, e), vv = ld(M.e2v, e);
  const uint32_t c1 = cc.x, c2 = cc.y;
  const int n = ld(M.nEoE, e);
  const double dc = ld(M.dc, e), dv = ld(M.dv, e);
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t o = e * L + k;
    const double pve = ld(R.pve, o), dpv = ld(dpve, o);
    double dq = 0.0;
    for (int j = 0; j < n; j++) {
      const uint32_t oj = uint32_t(ld(M.eoe, j * E + e)) * L + k;
      const double w = ld(M.woe, j * E + e);
      const double uj = ld(R.u, oj), pvj = ld(R.pve, oj), hej = ld(R.he, oj);
      dq += w * (ld(du, oj) * 0.5 * (pve + pvj) * hej + uj * 0.5 * (dpv + ld(dpve, oj)) * hej +
                 uj * 0.5 * (pve + pvj) * ld(dhe, oj));
    }
    double dtu = dq - (ld(dke, c2 * L + k) - ld(dke, c1 * L + k) +
                       M.gravity * (ld(dh, c2 * L + k) - ld(dh, c1 * L + k))) /
                          dc;
    if (P.use_visc)
      dtu += P.visc_coef * ((ld(ddiv, c2 * L + k) - ld(ddiv, c1 * L + k)) / dc -
                            (ld(dvort, uint32_t(vv.y) * L + k) -
                             ld(dvort, uint32_t(vv.x) * L + k)) / dv);
    const double d0 = ld(du0, o);
    if (P.s == 0) {
      st(dnu, o, d0 + P.w_rk * dtu);
      st(duy, o, d0 + P.w_sub * dtu);
    } else if (P.s < 3) {
      st(dnu, o, ld(dnu, o) + P.w_rk * dtu);
      st(duy, o, d0 + P.w_sub * dtu);
    } else {
      st(duy, o, ld(dnu, o) + P.w_rk * dtu);
    }
  }
}

// ---- adjoint ------------------------------------------------------------------

// Start of a step's reverse sweep from the adjoint of its final state:
// the final division T = Q / h reversed (aw_h = h adjoint after it, aw_tr =
// Q adjoint), the adjoint of the step's initial state seeded with the
// direct paths, and the stage-3 tendency adjoints.
// The prep of one edge level (o) from ub, the u adjoint of the step's
// final state, and of one cell level (ol) from hb and trb[t], its h and
// tracer adjoints; w_rk is the stage-3 weight.  k_ad_prep loads these
// values, or the last kernels of the following step's sweep (k_ad_cell and
// k_ad_edge_b at stage 0, which form them) call the same code.
struct PrepOut {
  double *ub_out, *hb_out, *trb_out, *aw_u, *aw_h, *aw_tr, *tb_u, *tb_h, *tb_tr;
};
__device__ __forceinline__ void prep_edge(uint32_t o, double ub, bool freeze_u, double w_rk,
                                          const PrepOut &Q) {
  st(Q.aw_u, o, ub);
  st(Q.ub_out, o, ub);
  st(Q.tb_u, o, freeze_u ? 0.0 : w_rk * ub);
}
__device__ __forceinline__ void prep_cell(uint32_t ol, uint32_t NT, double hb,
                                          const double (&trb)[8], const Rec &Rf, const Rec &R0,
                                          double w_rk, const PrepOut &Q) {
  const double hf = ld(Rf.h, ol), h0 = ld(R0.h, ol);
  double htot = hb;
  for (uint32_t t = 0; t < NT; t++) {
    const uint32_t ot = ol * NT + t;
    htot -= trb[t] * ld(Rf.tr, ot) / hf;
  }
  double xh = htot;
  for (uint32_t t = 0; t < NT; t++) {
    const uint32_t ot = ol * NT + t;
    const double qb = trb[t] / hf;
    st(Q.aw_tr, ot, qb);
    xh += qb * ld(R0.tr, ot);
    st(Q.trb_out, ot, qb * h0);
    st(Q.tb_tr, ot, w_rk * qb);
  }
  st(Q.aw_h, ol, htot);
  st(Q.hb_out, ol, xh);
  st(Q.tb_h, ol, w_rk * htot);
}

__global__ __launch_bounds__(kAdjMaxBlock) void k_ad_prep(DMesh M, AdjParams P, Rec Rf, Rec R0,
                                                const double *__restrict__ ub_in,
                                                const double *__restrict__ hb_in,
                                                const double *__restrict__ trb_in, PrepOut Q) {
  uint32_t k0, k1;
  int lane_;
  const int i = level_item(M.nL, k0, k1, lane_);
  const uint32_t L = M.nL, NT = P.nT;
  if (i < M.nE) {
    for (uint32_t k = k0; k < k1; k++) {
      const uint32_t o = i * L + k;
      prep_edge(o, ld(ub_in, o), P.freeze_u, P.w_rk, Q);
    }
  }
  if (i < M.nC) {
    for (uint32_t k = k0; k < k1; k++) {
      const uint32_t ol = i * L + k;
      double trb[8];
      for (uint32_t t = 0; t < NT; t++) trb[t] = ld(trb_in, ol * NT + t);
      prep_cell(ol, NT, ld(hb_in, ol), trb, Rf, R0, P.w_rk, Q);
    }
  }
}

// The next step's prep, fused into the stage-0 kernels of a step's sweep:
// on when `on`, with the records of the next step's start (R0n; its final
// state is this step's start) and the stage-3 weight.
struct PrepNext {
  int on;
  Rec R0n;
  double w_rk;
  PrepOut Q;
};

// Ragged lists are read in chunks of kChunk slots: the slot indices of a
// chunk are clamped to the row (the tail repeats the last slot, masked out
// of the sums), so every load of a chunk is in bounds and issues before the
// first value is needed, instead of one dependent round trip per slot.
constexpr int kChunk = 16;
// the advection stencil has at most 16 slots (maxEdges2): one chunk
static_assert(kChunk >= 16, "the edge part reads a cell's advection slots in one chunk");

// Edge part of the stage adjoint: from the tendency adjoints to the partial
// u adjoint (thickness flux, PV flux, tracer flux paths), the h_edge and
// pv_edge adjoints, and per (edge, tracer) what the cells gather: r1 / r2
// (into cell1 / cell2: upwind flux and mixing) and, per (stencil slot,
// edge, tracer), hbf = the 3rd-order flux adjoint times F times the slot's
// coefficient: the adjoint the slot's cell receives.  These three keep
// levels innermost ([edge][tracer][level], hbf [slot][edge][tracer][level]):
// the lanes of an edge's levels then store whole lines (with tracers
// innermost every store filled half of each line it touched).
// Two lanes per edge (adjacent, one wave) run the same code on alternate
// slots (or slot pairs) of every ragged list (edgesOnEdge, its transpose,
// the advection stencil); one lane exchange sums the partials and lane 0
// finishes.
constexpr int kEdgeLanes = 2;
// Levels from which the edge and cell parts run one lane per entity and
// level (the lanes of an entity's levels then cover whole lines).
constexpr int kOneLaneLevels = 8;
// The sums' exchange trees over kEdgeLanes (2) or kCellLanes (4) partials,
// v[m] being the partial of lane `lane + m * PL` of the PL physical lanes:
// with PL below the partial count one lane holds several partials (levels
// on lanes: one lane per entity and level keeps every load a contiguous run
// of levels) and adds them in the order of the exchange tree, so every
// launch shape gives the same bits.
template <int PL>
__device__ __forceinline__ double edge_lanes_sum(const double (&v)[kEdgeLanes / PL]) {
  if constexpr (PL == kEdgeLanes) return v[0] + __shfl_xor(v[0], 1);
  else return v[0] + v[1];
}
// FAST (meshes of at most hexagons: 10 padded slots, 5 slot pairs, with
// edgesOnEdge symmetric and advCoef_3rd in compact form): the edge part
// reads the forward kernels' slot-paired stencils, partial m taking the
// slot pairs m, m + kEdgeLanes, ...; otherwise the slot-major rows, partial
// m taking the slots m, m + kEdgeLanes, ...
constexpr int kNarrowPairs = 5;
constexpr int kFastPairs = (kNarrowPairs + kEdgeLanes - 1) / kEdgeLanes;  // per partial
template <bool FAST>
__device__ __forceinline__ int edge_slot(int vl, int i) {
  if constexpr (FAST) return 2 * (vl + (i / 2) * kEdgeLanes) + (i % 2);
  else return vl + i * kEdgeLanes;
}
template <int PL, bool FAST, bool UNI = false>
__global__ __launch_bounds__(kAdjMaxBlock) void k_ad_edge_a(DMesh M, AdjMesh A, AdjParams P, Rec R,
                                                  const double *__restrict__ tb_u,
                                                  const double *__restrict__ tb_h,
                                                  const double *__restrict__ tb_tr, double *ubA,
                                                  double *heb, double *pveb, double *r1,
                                                  double *r2, double *hbf) {
  static_assert(PL == 1 || PL == kEdgeLanes, "one lane or kEdgeLanes lanes per edge and level");
  constexpr int VPL = kEdgeLanes / PL;  // partials per lane
  constexpr int Q = kChunk / kEdgeLanes;  // slots per partial and round
  constexpr int NS = FAST ? 2 * kFastPairs : Q;  // advection slots per partial
  uint32_t k0, k1;
  int lane;
  static_assert(!UNI || PL == 1, "one edge per wave with levels on lanes only");
  int e = level_item<PL>(M.nL, k0, k1, lane);
  if constexpr (UNI) e = __builtin_amdgcn_readfirstlane(e);
  if (e >= M.nE) return;
  const uint32_t L = M.nL, E = M.nE, NT = P.nT;
  const int2 cc = ldm<UNI>(M.e2c, e);
  const uint32_t c1 = cc.x, c2 = cc.y;
  const double A1 = ldm<UNI>(M.area, c1), A2 = ldm<UNI>(M.area, c2);
  const double dv = ldm<UNI>(M.dv, e), dc = ldm<UNI>(M.dc, e);
  const int na = P.fct ? ldm<UNI>(M.nAdv, e) : 0;
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t o = e * L + k;
    const double ue = ld(R.u, o), hee = ld(R.he, o), pvee = ld(R.pve, o);
    double pb = 0.0, s1 = 0.0, s2 = 0.0;
    if (!P.freeze_u) {
      const int n = ldm<UNI>(M.nEoE, e);
      const double tue = ld(tb_u, o);
      double pbv[VPL], s1v[VPL], s2v[VPL];
      if constexpr (FAST) {
        // edgesOnEdge is symmetric: the transposed row holds the same edges,
        // e2 = edgesOnEdge(e, j) weighted by weightsOnEdge(e2, r), r the
        // slot of e in e2's row (nibble j of rslot), gathered from the
        // slot-paired weights that the edges near e stream themselves
        const uint64_t rv = ldm<UNI>(A.rslot, e);
#pragma unroll
        for (int m = 0; m < VPL; m++) {
          const int vl = lane + m * PL;
          uint32_t oj[2 * kFastPairs];
          double w[2 * kFastPairs], wr[2 * kFastPairs];
#pragma unroll
          for (int q = 0; q < kFastPairs; q++) {
            const int p = min(vl + q * kEdgeLanes, kNarrowPairs - 1);
            const mpas_i2v ij = ldmv<UNI, mpas_i2v>(M.eoeP, p * E + e);
            const mpas_d2v wv = ldmv<UNI, mpas_d2v>(M.woeP, p * E + e);
            oj[2 * q] = uint32_t(ij.x);
            oj[2 * q + 1] = uint32_t(ij.y);
            w[2 * q] = wv.x;
            w[2 * q + 1] = wv.y;
          }
#pragma unroll
          for (int i = 0; i < 2 * kFastPairs; i++) {
            const int j = edge_slot<true>(vl, i);
            const uint32_t r = uint32_t(rv >> (4 * j)) & 15u;
            wr[i] = ldm<UNI>(M.woeP, ((r / 2) * E + oj[i]) * 2 + (r % 2));
          }
          double pbx = 0.0, s1x = 0.0, s2x = 0.0;
#pragma unroll
          for (int i = 0; i < 2 * kFastPairs; i++) {
            const bool live = edge_slot<true>(vl, i) < n;
            const uint32_t oe = oj[i] * L + k;
            const double x = 0.5 * w[i] * tue * ld(R.u, oe) * ld(R.he, oe);
            pbx += live ? x : 0.0;
            const double t = ld(tb_u, oe) * wr[i];
            s1x += live ? t * 0.5 * (ld(R.pve, oe) + pvee) : 0.0;
            s2x += live ? t : 0.0;
          }
          pbv[m] = pbx;
          s1v[m] = s1x;
          s2v[m] = s2x;
        }
      } else {
        const int nt = ld(A.nEt, e);
#pragma unroll
        for (int m = 0; m < VPL; m++) {
          const int vl = lane + m * PL;
          double pbx = 0.0, s1x = 0.0, s2x = 0.0;
          for (int j0 = vl; j0 < n; j0 += kChunk) {
            uint32_t oj[Q];
            double w[Q];
#pragma unroll
            for (int j = 0; j < Q; j++) {
              const int jj = min(j0 + j * kEdgeLanes, n - 1);
              oj[j] = uint32_t(ld(M.eoe, jj * E + e)) * L + k;
              w[j] = ld(M.woe, jj * E + e);
            }
#pragma unroll
            for (int j = 0; j < Q; j++) {
              const double x = 0.5 * w[j] * tue * ld(R.u, oj[j]) * ld(R.he, oj[j]);
              pbx += (j0 + j * kEdgeLanes < n) ? x : 0.0;
            }
          }
          for (int j0 = vl; j0 < nt; j0 += kChunk) {
            uint32_t o2[Q];
            double w[Q];
#pragma unroll
            for (int j = 0; j < Q; j++) {
              const int jj = min(j0 + j * kEdgeLanes, nt - 1);
              o2[j] = uint32_t(ld(A.et, jj * E + e)) * L + k;
              w[j] = ld(A.wt, jj * E + e);
            }
#pragma unroll
            for (int j = 0; j < Q; j++) {
              const double t = ld(tb_u, o2[j]) * w[j];
              const bool live = j0 + j * kEdgeLanes < nt;
              s1x += live ? t * 0.5 * (ld(R.pve, o2[j]) + pvee) : 0.0;
              s2x += live ? t : 0.0;
            }
          }
          pbv[m] = pbx;
          s1v[m] = s1x;
          s2v[m] = s2x;
        }
      }
      pb = edge_lanes_sum<PL>(pbv);
      s1 = edge_lanes_sum<PL>(s1v);
      s2 = edge_lanes_sum<PL>(s2v);
    }
    // tracer paths: the stencil sums split over the partials, the rest once
    double Fb = 0.0, hbt = 0.0;
    if (P.fct) {
      const double rho = M.msd * P.kappa * dv / dc;
      const double F = ue * hee, sgn = fsign(F);
      const double Fp = F > 0.0 ? F : 0.0, Fm = F < 0.0 ? F : 0.0;
      // the partials' advection slots (every slot of the row: one chunk),
      // loaded once for every tracer
      uint32_t ib[VPL][NS];
      double cf[VPL][NS];
      if constexpr (FAST) {
        // advCoef_3rd from its 2-bit side code: +-advCoef, or the edge's
        // explicit value numbered by the explicit slots before it
        const uint32_t side = ldm<UNI>(M.advSide, e);
        const uint32_t expl = ~(side | (side >> 1)) & 0x55555555u;  // bit 2i: slot i explicit
        double x[kAdvExplicit];
#pragma unroll
        for (int p = 0; p < kAdvExplicit / 2; p++) {
          const mpas_d2v v = ldmv<UNI, mpas_d2v>(M.advC3xP, p * E + e);
          x[2 * p] = v.x;
          x[2 * p + 1] = v.y;
        }
#pragma unroll
        for (int m = 0; m < VPL; m++) {
          const int vl = lane + m * PL;
#pragma unroll
          for (int q = 0; q < kFastPairs; q++) {
            const int p = min(vl + q * kEdgeLanes, kNarrowPairs - 1);
            const mpas_i2v ij = ldmv<UNI, mpas_i2v>(M.advP, p * E + e);
            const mpas_d2v a1 = ldmv<UNI, mpas_d2v>(M.advCoefP, p * E + e);
#pragma unroll
            for (int h = 0; h < 2; h++) {
              const int i = 2 * p + h;  // the (clamped) slot
              const double c = h ? a1.y : a1.x;
              const uint32_t code = (side >> (2 * i)) & 3u;
              const int nx = __popc(expl & ((1u << (2 * i)) - 1u));
              double xe = x[kAdvExplicit - 1];
#pragma unroll
              for (int z = kAdvExplicit - 2; z >= 0; z--) xe = nx == z ? x[z] : xe;
              const double a3 = code == 1u ? c : (code == 2u ? -c : xe);
              ib[m][2 * q + h] = uint32_t(h ? ij.y : ij.x) * L + k;
              cf[m][2 * q + h] = c + a3 * M.coef3rd * sgn;
            }
          }
        }
      } else {
#pragma unroll
        for (int m = 0; m < VPL; m++)
#pragma unroll
          for (int j = 0; j < NS; j++) {
            const int i = max(min(edge_slot<false>(lane + m * PL, j), na - 1), 0);
            ib[m][j] = uint32_t(ld(M.adv, i * E + e)) * L + k;
            cf[m][j] = ld(M.advCoef, i * E + e) + ld(M.advCoef3, i * E + e) * M.coef3rd * sgn;
          }
      }
      // the finishing lane's part of tracer t: the upwind and mixing paths
      auto finish = [&](uint32_t t, double phib, double lob, double hob, double hs, double T1,
                        double T2) {
        Fb += lob * dv * ((F > 0.0 ? T1 : 0.0) + (F < 0.0 ? T2 : 0.0)) + hob * hs;
        double a1 = lob * dv * Fp, a2 = lob * dv * Fm;
        if (P.use_hmix) {  // mixing flux h_edge (T2 - T1) rho enters with -sign/area
          const double Hb = -phib;
          hbt += Hb * rho * (T2 - T1);
          const double Db = Hb * hee * rho;
          a1 -= Db;
          a2 += Db;
        }
        st(r1, (e * NT + t) * L + k, a1);
        st(r2, (e * NT + t) * L + k, a2);
      };
      if (PL == 1 && NT % 2 == 0) {
        // levels on lanes: tracers two at a time, each pair's tracers,
        // limiter factors, tendency adjoints and antidiffusive fluxes
        // adjacent in their records and read with 16-byte accesses; each
        // tracer's arithmetic and order as one at a time (with two lanes
        // per edge the second tracer's registers would cost a wave per
        // SIMD)
        for (uint32_t t = 0; t < NT; t += 2) {
          const uint32_t i1 = (c1 * L + k) * NT + t, i2 = (c2 * L + k) * NT + t;
          const mpas_d2v b1 = ld_vec<false, mpas_d2v>(tb_tr, i1 / 2);
          const mpas_d2v b2 = ld_vec<false, mpas_d2v>(tb_tr, i2 / 2);
          const mpas_d2v f1a = ld_vec<false, mpas_d2v>(R.rfac, i1);
          const mpas_d2v f1b = ld_vec<false, mpas_d2v>(R.rfac, i1 + 1);
          const mpas_d2v f2a = ld_vec<false, mpas_d2v>(R.rfac, i2);
          const mpas_d2v f2b = ld_vec<false, mpas_d2v>(R.rfac, i2 + 1);
          const mpas_d2v an = ld_vec<false, mpas_d2v>(R.anti, (o * NT + t) / 2);
          const double phib[2] = {-b1.x / A1 + b2.x / A2, -b1.y / A1 + b2.y / A2};
          const double lam[2] = {slope_of(f1a.x, f1a.y, f2a.x, f2a.y, an.x),
                                 slope_of(f1b.x, f1b.y, f2b.x, f2b.y, an.y)};
          double lob[2], hob[2], hf[2];
#pragma unroll
          for (int q = 0; q < 2; q++) {
            lob[q] = (1.0 - lam[q]) * phib[q];
            hob[q] = lam[q] * phib[q];
            hf[q] = hob[q] * F;
          }
          double hsv0[VPL], hsv1[VPL];
#pragma unroll
          for (int m = 0; m < VPL; m++) {
            double h0 = 0.0, h1 = 0.0;
#pragma unroll
            for (int j = 0; j < NS; j++) {
              const int i = edge_slot<FAST>(lane + m * PL, j);
              if (i < na) {
                const mpas_d2v tv = ld_vec<false, mpas_d2v>(R.tr, (ib[m][j] * NT + t) / 2);
                h0 += tv.x * cf[m][j];
                h1 += tv.y * cf[m][j];
                double *const hs = hbf + uint64_t(i) * A.hslot;
                st(hs, (e * NT + t) * L + k, hf[0] * cf[m][j]);
                st(hs, (e * NT + t + 1) * L + k, hf[1] * cf[m][j]);
              }
            }
            hsv0[m] = h0;
            hsv1[m] = h1;
          }
          const double hs0 = edge_lanes_sum<PL>(hsv0), hs1 = edge_lanes_sum<PL>(hsv1);
          if (lane) continue;
          const mpas_d2v T1 = ld_vec<false, mpas_d2v>(R.tr, i1 / 2);
          const mpas_d2v T2 = ld_vec<false, mpas_d2v>(R.tr, i2 / 2);
          finish(t, phib[0], lob[0], hob[0], hs0, T1.x, T2.x);
          finish(t + 1, phib[1], lob[1], hob[1], hs1, T1.y, T2.y);
        }
      } else {
        for (uint32_t t = 0; t < NT; t++) {
          const uint32_t i1 = (c1 * L + k) * NT + t, i2 = (c2 * L + k) * NT + t;
          const double phib = -ld(tb_tr, i1) / A1 + ld(tb_tr, i2) / A2;
          const double lam = frozen_slope(R.rfac, i1, i2, ld(R.anti, o * NT + t));
          const double lob = (1.0 - lam) * phib, hob = lam * phib;
          const double hf = hob * F;
          double hsv[VPL];
#pragma unroll
          for (int m = 0; m < VPL; m++) {
            double hsx = 0.0;
#pragma unroll
            for (int j = 0; j < NS; j++) {
              const int i = edge_slot<FAST>(lane + m * PL, j);
              if (i < na) {
                hsx += ld(R.tr, ib[m][j] * NT + t) * cf[m][j];
                st(hbf + uint64_t(i) * A.hslot, (e * NT + t) * L + k, hf * cf[m][j]);
              }
            }
            hsv[m] = hsx;
          }
          const double hs = edge_lanes_sum<PL>(hsv);
          if (lane) continue;
          finish(t, phib, lob, hob, hs, ld(R.tr, i1), ld(R.tr, i2));
        }
      }
    }
    if (lane) continue;
    // tend_h: cell1 -= u dv h_edge / area, cell2 += (solve.cc:191-205)
    const double fb = -ld(tb_h, c1 * L + k) / A1 + ld(tb_h, c2 * L + k) / A2;
    const double ub = fb * dv * hee + s1 * hee + Fb * hee;
    const double hb = fb * ue * dv + s1 * ue + hbt + Fb * ue;
    st(ubA, o, ub);
    st(heb, o, hb);
    st(pveb, o, pb + 0.5 * s2 * ue * hee);
  }
}

// Vertex part: pv_vertex = (f + vorticity) / h_vertex and the viscosity's
// vorticity term reversed; vT and hvT are the vorticity and h_vertex
// adjoints divided by areaTriangle.  vertex_pvvb forms the pv_vertex
// adjoint and the recorded h_vertex and pv_vertex of one vertex level; the
// cell part, merged with the vertex part into one launch, forms hvT of its
// vertices itself with the same code (the same bits).
struct VertexAdj {
  int x[3];
  double pvvb, hv, pvv, at;
};
__device__ __forceinline__ VertexAdj vertex_pvvb(const DMesh &M, const Rec &R,
                                                 const double *__restrict__ pveb, int v,
                                                 uint32_t k) {
  const uint32_t L = M.nL, V = M.nV;
  VertexAdj a;
  uint32_t c[3];
  double kt[3];
  for (int i = 0; i < 3; i++) {
    a.x[i] = ld(M.v2e, i * V + v);
    c[i] = ld(M.v2c, i * V + v);
    kt[i] = ld(M.kite, i * V + v);
  }
  a.at = ld(M.areaTri, v);
  a.pvvb = 0.0;
  for (int i = 0; i < 3; i++) a.pvvb += 0.5 * ld(pveb, v2e_id(a.x[i]) * L + k);
  vertex_state(M, v, k, R, c, kt, a.at, a.hv, a.pvv);
  return a;
}
__device__ __forceinline__ double vertex_hvT(const VertexAdj &a) {
  return -a.pvvb * a.pvv / a.hv / a.at;
}
// HV: write hvT too (the separate vertex launch; the merged launch's cells
// form it themselves)
template <bool HV>
__device__ __forceinline__ void ad_vertex_body(const DMesh &M, const AdjParams &P, const Rec &R,
                                               const double *__restrict__ pveb,
                                               const double *__restrict__ tb_u, double *vT,
                                               double *hvT, int item) {
  uint32_t k0, k1;
  int lane_;
  const int v = level_item_at(item, M.nL, k0, k1, lane_);
  if (v >= M.nV) return;
  const uint32_t L = M.nL;
  for (uint32_t k = k0; k < k1; k++) {
    const VertexAdj a = vertex_pvvb(M, R, pveb, v, k);
    double vb = a.pvvb / a.hv;
    if (P.use_visc && !P.freeze_u)
      for (int i = 0; i < 3; i++) {
        const uint32_t e = v2e_id(a.x[i]);
        vb += (-v2e_sign(a.x[i])) * P.visc_coef * ld(tb_u, e * L + k) / ld(M.dv, e);
      }
    st(vT, v * L + k, vb / a.at);
    if (HV) st(hvT, v * L + k, vertex_hvT(a));
  }
}
__global__ __launch_bounds__(kAdjMaxBlock) void k_ad_vertex(DMesh M, AdjParams P, Rec R,
                                                  const double *__restrict__ pveb,
                                                  const double *__restrict__ tb_u, double *vT,
                                                  double *hvT) {
  ad_vertex_body<true>(M, P, R, pveb, tb_u, vT, hvT, xcd_item());
}

// Cell part: h and tracer adjoints of the stage input, ke and divergence
// adjoints for the edges, then the bookkeeping: X_s of stages 1-3 was
// formed from the step's initial state and the stage s-1 tendencies.
// Four lanes per cell (adjacent, one wave): lane r takes every 4th entry
// of the cell's edges, vertices and stencil slots; a fixed exchange tree
// sums the partials and lane 0 finishes.
constexpr int kCellLanes = 4;
// edges of a cell (at most octagons), and per partial
constexpr int kMaxCellSlots = 8;
constexpr int kCellSlots = kMaxCellSlots / kCellLanes;
// advection pairs per partial read in one straight-line chunk (a cell of a
// hexagon mesh sits in about 30 edges' stencils), and the tracers whose
// own values the finishing lane has loaded with the gathers
constexpr int kAtSlots = 8;
constexpr int kHoistT = 2;
__device__ __forceinline__ double cell_lanes_sum(double x) {
  x += __shfl_xor(x, 1);
  x += __shfl_xor(x, 2);
  return x;
}
// The same tree over the partials v[m] of lanes lane + m * PL (edge_lanes_sum).
template <int PL>
__device__ __forceinline__ double cell_lanes_sum(const double (&v)[kCellLanes / PL]) {
  if constexpr (PL == kCellLanes) return cell_lanes_sum(v[0]);
  else return (v[0] + v[1]) + (v[2] + v[3]);
}
// The cell part's arguments (one struct: the merged launch passes them on).
struct CellArgs {
  const double *tb_u, *heb, *hvT, *pveb, *r1, *r2, *hbf;
  double *keT, *divT, *xb_h, *xb_tr;
  const double *aw_h, *aw_tr;
  double *tb_h, *tb_tr;
};
// VF: the h_vertex adjoints of the cell's vertices formed here from pveb
// (vertex_pvvb), not read from hvT
template <int PL, bool VF, bool UNI = false>
__device__ __forceinline__ void ad_cell_body(const DMesh &M, const AdjMesh &A, const AdjParams &P,
                                             const Rec &R, const Rec &R0, const CellArgs &G,
                                             const PrepNext &N, int item) {
  static_assert(PL == 1 || PL == kCellLanes, "one lane or kCellLanes lanes per cell and level");
  constexpr int VPL = kCellLanes / PL;  // partials per lane
  const double *__restrict__ tb_u = G.tb_u;
  const double *__restrict__ heb = G.heb;
  const double *__restrict__ hvT = G.hvT;
  const double *__restrict__ r1 = G.r1;
  const double *__restrict__ r2 = G.r2;
  const double *__restrict__ hbf = G.hbf;
  double *keT = G.keT, *divT = G.divT, *xb_h = G.xb_h, *xb_tr = G.xb_tr;
  const double *aw_h = G.aw_h, *aw_tr = G.aw_tr;
  double *tb_h = G.tb_h, *tb_tr = G.tb_tr;
  uint32_t k0, k1;
  int lane;
  static_assert(!UNI || PL == 1, "one cell per wave with levels on lanes only");
  int c = level_item_at<PL>(item, M.nL, k0, k1, lane);
  if constexpr (UNI) c = __builtin_amdgcn_readfirstlane(c);  // one cell per wave
  if (c >= M.nC) return;
  const uint32_t L = M.nL, C = M.nC, NT = P.nT;
  const int n = ldm<UNI>(M.nEoC, c), nv = ldm<UNI>(A.nVc, c);
  const int na = P.fct ? ldm<UNI>(A.nAt, c) : 0;
  const uint8_t neg = ldm<UNI>(M.sgnNeg, c);
  const double ar = ldm<UNI>(M.area, c);
  constexpr int Q = kChunk / kCellLanes;  // entries per partial and round of the tail
  // row slots clamped into the padded rows (uniform widths, so the id
  // loads need not wait for the row counts; the padding holds valid ids)
  const int wc = M.narrow ? 6 : 8;
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t ol = c * L + k;
    double g1v[VPL], hsumv[VPL], Tbv[VPL][8];
#pragma unroll
    for (int m = 0; m < VPL; m++) {
      const int vl = lane + m * PL;
      double g1 = 0.0, hsum = 0.0, Tb[8] = {};
      // the first kCellSlots edges and vertices and kAtSlots advection
      // pairs of the partial's share read in straight-line code, the rest
      // of longer rows after
      uint32_t ej[kCellSlots], vj[kCellSlots], aj[kAtSlots];
      double kv[kCellSlots];
#pragma unroll
      for (int q = 0; q < kCellSlots; q++) {
        const int j = vl + q * kCellLanes;
        ej[q] = ldm<UNI>(M.c2e, min(j, wc - 1) * C + c);
        const int i = min(j, A.wVc - 1);
        vj[q] = ldm<UNI>(A.vc, i * C + c);
        kv[q] = ldm<UNI>(A.vk, i * C + c);
      }
      if (P.fct)
#pragma unroll
        for (int q = 0; q < kAtSlots; q++)
          aj[q] = uint32_t(ldm<UNI>(A.at, min(vl + q * kCellLanes, A.wAt - 1) * C + c));
#pragma unroll
      for (int q = 0; q < kCellSlots; q++) {
        const int j = vl + q * kCellLanes;
        const bool live = j < n;
        const uint32_t e = ej[q], oe = e * L + k;
        if (!P.freeze_u) g1 += live ? sign_on_cell(neg, j) * ld(tb_u, oe) / ldm<UNI>(M.dc, e) : 0.0;
        hsum += live ? 0.5 * ld(heb, oe) : 0.0;
        const bool first = (neg >> j) & 1;
        for (uint32_t t = 0; t < NT; t++)
          Tb[t] += live ? (first ? ld(r1, (e * NT + t) * L + k) : ld(r2, (e * NT + t) * L + k))
                        : 0.0;
      }
#pragma unroll
      for (int q = 0; q < kCellSlots; q++) {
        const double hv =
            VF ? vertex_hvT(vertex_pvvb(M, R, G.pveb, vj[q], k)) : ld(hvT, vj[q] * L + k);
        hsum += vl + q * kCellLanes < nv ? hv * kv[q] : 0.0;
      }
      for (int i = vl + kCellSlots * kCellLanes; i < nv; i += kCellLanes) {
        const uint32_t v = ld(A.vc, i * C + c);
        const double hv = VF ? vertex_hvT(vertex_pvvb(M, R, G.pveb, v, k)) : ld(hvT, v * L + k);
        hsum += hv * ld(A.vk, i * C + c);
      }
      if (P.fct) {
        for (uint32_t t = 0; t < NT; t++) {
#pragma unroll
          for (int q = 0; q < kAtSlots; q++)
            Tb[t] += (vl + q * kCellLanes < na) ? ld(hbf_at(A, hbf, aj[q]), (at_edge(aj[q]) * NT + t) * L + k) : 0.0;
        }
        for (int i0 = vl + kAtSlots * kCellLanes; i0 < na; i0 += kChunk) {
          uint32_t os[Q];
#pragma unroll
          for (int j = 0; j < Q; j++)
            os[j] = uint32_t(ldm<UNI>(A.at, min(i0 + j * kCellLanes, na - 1) * C + c));
          for (uint32_t t = 0; t < NT; t++) {
#pragma unroll
            for (int j = 0; j < Q; j++)
              Tb[t] += (i0 + j * kCellLanes < na) ? ld(hbf_at(A, hbf, os[j]), (at_edge(os[j]) * NT + t) * L + k) : 0.0;
          }
        }
      }
      g1v[m] = g1;
      hsumv[m] = hsum;
#pragma unroll
      for (int t = 0; t < 8; t++) Tbv[m][t] = Tb[t];
    }
    // the finishing lane's own values (first kHoistT tracers), loaded by
    // every lane of the cell before the exchange: behind the branch on the
    // lane they would wait for the sums
    const bool late = P.s >= 1;
    const double xbh = ld(xb_h, ol);
    double hs = 0.0, h0 = 0.0, awh = 0.0;
    if (late) {
      hs = ld(R.h, ol);
      h0 = ld(R0.h, ol);
      awh = ld(aw_h, ol);
    }
    double xbt[kHoistT] = {}, trs[kHoistT] = {}, tr0[kHoistT] = {}, awt[kHoistT] = {};
#pragma unroll
    for (int t = 0; t < kHoistT; t++) {
      if (uint32_t(t) >= NT) break;  // uniform
      const uint32_t ot = ol * NT + t;
      xbt[t] = ld(xb_tr, ot);
      if (late) {
        trs[t] = ld(R.tr, ot);
        tr0[t] = ld(R0.tr, ot);
        awt[t] = ld(aw_tr, ot);
      }
    }
    const double g1 = cell_lanes_sum<PL>(g1v), hsum = cell_lanes_sum<PL>(hsumv);
    double Tb[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
      if (uint32_t(t) >= NT) break;  // uniform
      double x[VPL];
#pragma unroll
      for (int m = 0; m < VPL; m++) x[m] = Tbv[m][t];
      Tb[t] = cell_lanes_sum<PL>(x);
    }

    if (lane) continue;
    const double hb = -M.gravity * g1 + hsum;
    st(keT, ol, -g1 / ar);
    if (P.use_visc) st(divT, ol, P.visc_coef * g1 / ar);
    if (late) {
      double htot = hb;
      for (uint32_t t = 0; t < NT; t++)
        htot -= Tb[t] * (t < kHoistT ? trs[t] : ld(R.tr, ol * NT + t)) / hs;
      double xh = xbh + htot;
      for (uint32_t t = 0; t < NT; t++) {
        const uint32_t ot = ol * NT + t;
        const bool hx = t < kHoistT;
        const double pb = Tb[t] / hs;
        xh += pb * (hx ? tr0[t] : ld(R0.tr, ot));
        st(xb_tr, ot, (hx ? xbt[t] : ld(xb_tr, ot)) + pb * h0);
        st(tb_tr, ot, P.w_rk_m * (hx ? awt[t] : ld(aw_tr, ot)) + P.w_sub_m * pb);
      }
      st(xb_h, ol, xh);
      st(tb_h, ol, P.w_rk_m * awh + P.w_sub_m * htot);
    } else {
      // the adjoint of the step's initial state is complete here
      const double xh = xbh + hb;
      st(xb_h, ol, xh);
      double xt[8];
      for (uint32_t t = 0; t < NT; t++) {
        const uint32_t ot = ol * NT + t;
        xt[t] = (t < kHoistT ? xbt[t] : ld(xb_tr, ot)) + Tb[t];
        st(xb_tr, ot, xt[t]);
      }
      // the next step's prep (its final state is this step's start, R0)
      if (N.on) prep_cell(ol, NT, xh, xt, R0, N.R0n, N.w_rk, N.Q);
    }
  }
}

template <int PL, bool UNI = false>
__global__ __launch_bounds__(kAdjMaxBlock) void k_ad_cell(DMesh M, AdjMesh A, AdjParams P, Rec R,
                                                Rec R0, CellArgs G, PrepNext N) {
  ad_cell_body<PL, false, UNI>(M, A, P, R, R0, G, N, xcd_item());
}
// The vertex part (vT only: the workgroups before vblocks) and the cell
// part (forming hvT itself) in one launch.
template <int PL>
__global__ __launch_bounds__(kAdjMaxBlock) void k_ad_cell_vertex(DMesh M, AdjMesh A, AdjParams P,
                                                       Rec R, Rec R0, CellArgs G, PrepNext N,
                                                       double *vT, int vblocks) {
  const int b = blockIdx.x;
  if (b < vblocks)
    ad_vertex_body<false>(M, P, R, G.pveb, G.tb_u, vT, nullptr, xcd_item_in(b, vblocks));
  else
    ad_cell_body<PL, true>(M, A, P, R, R0, G, N,
                           xcd_item_in(b - vblocks, int(gridDim.x) - vblocks));
}

// Edge part 2: the u adjoint of the stage input completed with the paths
// through ke, the divergence and the vorticity, then the bookkeeping of u.
__global__ __launch_bounds__(kAdjMaxBlock) void k_ad_edge_b(DMesh M, AdjParams P, Rec R,
                                                  const double *__restrict__ ubA,
                                                  const double *__restrict__ keT,
                                                  const double *__restrict__ divT,
                                                  const double *__restrict__ vT, double *xb_u,
                                                  const double *aw_u, double *tb_u,
                                                  PrepNext N) {
  uint32_t k0, k1;
  int lane_;
  const int e = level_item(M.nL, k0, k1, lane_);
  if (e >= M.nE) return;
  const uint32_t L = M.nL;
  const int2 cc = ld(M.e2c, e), vv = ld(M.e2v, e);
  const uint32_t c1 = cc.x, c2 = cc.y;
  const double dc = ld(M.dc, e), dv = ld(M.dv, e);
  for (uint32_t k = k0; k < k1; k++) {
    const uint32_t o = e * L + k;
    double ub = ld(ubA, o) +
                0.5 * dc * dv * ld(R.u, o) * (ld(keT, c1 * L + k) + ld(keT, c2 * L + k));
    if (P.use_visc) ub += dv * (ld(divT, c1 * L + k) - ld(divT, c2 * L + k));
    ub += dc * (ld(vT, uint32_t(vv.y) * L + k) - ld(vT, uint32_t(vv.x) * L + k));
    const double xu = ld(xb_u, o) + ub;
    st(xb_u, o, xu);
    if (P.s >= 1) st(tb_u, o, P.freeze_u ? 0.0 : P.w_rk_m * ld(aw_u, o) + P.w_sub_m * ub);
    else if (N.on) prep_edge(o, xu, P.freeze_u, N.w_rk, N.Q);  // the next step's prep
  }
}

// ---- host side ----------------------------------------------------------------

int dalloc(AdjState &A, size_t n, double **out) {
  *out = nullptr;
  if (!n) return 0;
  // the kernels address every array with 32-bit byte offsets
  if (n > size_t(UINT32_MAX) / sizeof(double))
    MPAS_FAIL(MPAS_ESHAPE, "an adjoint tape array of " + std::to_string(sizeof(double) * n) +
                               " bytes exceeds the 4 GiB the kernels address");
  void *p = nullptr;
  if (hipMalloc(&p, sizeof(double) * n) != hipSuccess)
    MPAS_FAIL(MPAS_ENOMEM, "hipMalloc of " + std::to_string(sizeof(double) * n) +
                               " bytes for the adjoint tape failed");
  A.bufs.push_back(p);
  MPAS_HIP(zero_now(p, sizeof(double) * n));
  *out = static_cast<double *>(p);
  return 0;
}

// `parts` arrays of `part` doubles each in one allocation: each part stays
// within the 4 GiB one 32-bit offset addresses (the kernels address each
// from its own base), the whole may pass it.
int dalloc_parts(AdjState &A, size_t part, int parts, double **out) {
  *out = nullptr;
  if (!part || parts < 1) return 0;
  if (part > size_t(UINT32_MAX) / sizeof(double))
    MPAS_FAIL(MPAS_ESHAPE, "a part of an adjoint tape array of " +
                               std::to_string(sizeof(double) * part) +
                               " bytes exceeds the 4 GiB the kernels address");
  const size_t bytes = sizeof(double) * part * size_t(parts);
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess)
    MPAS_FAIL(MPAS_ENOMEM, "hipMalloc of " + std::to_string(bytes) +
                               " bytes for the adjoint tape failed");
  A.bufs.push_back(p);
  MPAS_HIP(zero_now(p, bytes));
  *out = static_cast<double *>(p);
  return 0;
}

template <class T>
int upload_vec(AdjState &A, const std::vector<T> &v, const T **out) {
  *out = nullptr;
  if (v.empty()) return 0;
  void *p = nullptr;
  if (hipMalloc(&p, sizeof(T) * v.size()) != hipSuccess)
    MPAS_FAIL(MPAS_ENOMEM, "hipMalloc for the transposed mesh failed");
  A.bufs.push_back(p);
  MPAS_HIP(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  *out = static_cast<const T *>(p);
  return 0;
}

// A ragged list of (id, one or two values) per entity as slot-major ELL.
struct Ragged {
  std::vector<std::vector<int>> id;
  std::vector<std::vector<double>> a, b;
  explicit Ragged(int n) : id(n), a(n), b(n) {}
  int width() const {
    size_t w = 1;
    for (auto &r : id) w = std::max(w, r.size());
    return int(w);
  }
};

int upload_ragged(AdjState &A, const Ragged &r, bool two, int *width, const uint8_t **cnt,
                  const int **ids, const double **va, const double **vb) {
  const int n = int(r.id.size()), w = r.width();
  if (w > 255) MPAS_FAIL(MPAS_ESHAPE, "transposed list wider than 255 entries");
  std::vector<uint8_t> c(n);
  std::vector<int> ix(size_t(w) * n, 0);
  std::vector<double> x(size_t(w) * n, 0.0), y(two ? size_t(w) * n : 0, 0.0);
  for (int i = 0; i < n; i++) {
    c[i] = uint8_t(r.id[i].size());
    for (size_t j = 0; j < r.id[i].size(); j++) {
      ix[j * n + i] = r.id[i][j];
      x[j * n + i] = r.a[i][j];
      if (two) y[j * n + i] = r.b[i][j];
    }
  }
  *width = w;
  int rc;
  if ((rc = upload_vec(A, c, cnt)) || (rc = upload_vec(A, ix, ids)) || (rc = upload_vec(A, x, va)))
    return rc;
  if (two && (rc = upload_vec(A, y, vb))) return rc;
  return 0;
}

int build_transposes(mpas_ctx *ctx, AdjState &A) {
  const HostMesh &m = ctx->host;
  const int C = m.nC, E = m.nE, V = m.nV;
  Ragged eT(E), aT(C), vC(C);
  for (int e2 = 0; e2 < E; e2++)
    for (int j = 0; j < m.e2e.size(e2); j++) {
      const int e = m.e2e.at(e2, j);
      eT.id[e].push_back(e2);
      eT.a[e].push_back(m.weightsOnEdge[size_t(e2) * 16 + j]);
    }
  if (m.fct && E >= (1 << kAtSlotShift))
    MPAS_FAIL(MPAS_ESHAPE, "the adjoint's advection transpose holds edge ids below 2^24 (" +
                               std::to_string(E) + " edges)");
  if (m.fct)
    for (int e = 0; e < E; e++)
      for (int i = 0; i < m.nAdvCellsForEdge[e]; i++) {
        const int c = m.advCellsForEdge[size_t(e) * 16 + i];
        aT.id[c].push_back(i << kAtSlotShift | e);
        aT.a[c].push_back(0.0);
      }
  for (int v = 0; v < V; v++)
    for (int i = 0; i < m.v2c.size(v); i++) {
      const int c = m.v2c.at(v, i);
      vC.id[c].push_back(v);
      vC.a[c].push_back(m.kite[size_t(v) * 3 + i]);
    }
  AdjMesh &T = A.T;
  int rc, w = 0;
  if ((rc = upload_ragged(A, eT, false, &w, &T.nEt, &T.et, &T.wt, nullptr))) return rc;
  // reverse slots, where every row of edgesOnEdge holds distinct edges and
  // its transposed row the same ones (then e sits in exactly one slot of
  // each row it appears in)
  bool sym = true;
  std::vector<uint64_t> rslot(E, 0);
  for (int e = 0; e < E && sym; e++) {
    const int n = m.e2e.size(e);
    std::vector<int> a(eT.id[e]), b;
    for (int j = 0; j < n; j++) b.push_back(m.e2e.at(e, j));
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    sym = n <= 16 && a == b && std::adjacent_find(b.begin(), b.end()) == b.end();
    for (int j = 0; j < n && sym; j++) {
      const int e2 = m.e2e.at(e, j);
      int r = -1;
      for (int i = 0; i < m.e2e.size(e2); i++)
        if (m.e2e.at(e2, i) == e) r = i;
      if (r < 0) sym = false;
      else rslot[e] |= uint64_t(r) << (4 * j);
    }
  }
  const DMesh &M = ctx->dm;
  const char *pe = std::getenv("MPAS_HIP_ADJ_PAIRED");
  A.edge_fast = sym && M.narrow && M.eoeP && M.woeP &&
                (!m.fct || (M.advP && M.advCoefP && M.advSide && M.advC3xP)) &&
                !(pe && pe[0] == '0');
  if (A.edge_fast && (rc = upload_vec(A, rslot, &T.rslot))) return rc;
  if ((rc = upload_ragged(A, vC, false, &w, &T.nVc, &T.vc, &T.vk, nullptr))) return rc;
  T.wVc = w;
  const double *unused = nullptr;
  T.wAt = 1;
  if (m.fct && (rc = upload_ragged(A, aT, false, &T.wAt, &T.nAt, &T.at, &unused, nullptr)))
    return rc;
  return 0;
}

int alloc_rec(AdjState &A, const DMesh &M, Rec &r) {
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, V = size_t(M.nV) * M.nL;
  const size_t T = C * M.nT;
  int rc;
  if ((rc = dalloc(A, E, &r.u)) || (rc = dalloc(A, C, &r.h)) || (rc = dalloc(A, T, &r.tr)) ||
      (rc = dalloc(A, E, &r.he)) || (rc = dalloc(A, C, &r.ke)) || (rc = dalloc(A, E, &r.pve)) ||
      (rc = dalloc(A, V, &r.vort)) || (rc = dalloc(A, C, &r.div)) ||
      (rc = dalloc(A, 2 * T, &r.rfac)) || (rc = dalloc(A, E * M.nT, &r.anti)))
    return rc;
  return 0;
}

int adj_alloc(mpas_ctx *ctx, int K) {
  const DMesh &M = ctx->dm;
  auto *A = new AdjState();
  const char *fp = std::getenv("MPAS_HIP_ADJ_FUSE_PREP");
  A->fuse_prep = !(fp && fp[0] == '0');
  const char *un = std::getenv("MPAS_HIP_ADJ_UNIFORM");
  A->uni_ok = !(un && un[0] == '0');
  const char *mv = std::getenv("MPAS_HIP_ADJ_MERGE_VERTEX");
  A->merge_vertex = !(mv && mv[0] == '0');
  const char *gp = std::getenv("MPAS_HIP_ADJ_GRAPH");
  A->graph_ok = !(gp && gp[0] == '0');
  ctx->adj = A;
  A->K = K;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, V = size_t(M.nV) * M.nL;
  const size_t T = C * M.nT, ET = E * M.nT;
  int rc;
  if ((rc = build_transposes(ctx, *A))) return rc;
  A->rec.resize(size_t(4) * K + 1);
  for (Rec &r : A->rec)
    if ((rc = alloc_rec(*A, M, r))) return rc;
  double **e_bufs[] = {&A->nu, &A->dnu, &A->dhe, &A->dpve, &A->aw_u, &A->tb_u, &A->ubA,
                       &A->heb, &A->pveb, &A->xb_u[0], &A->xb_u[1]};
  double **c_bufs[] = {&A->nh, &A->dnh, &A->dke, &A->ddiv, &A->aw_h, &A->tb_h, &A->keT,
                       &A->divT, &A->xb_h[0], &A->xb_h[1]};
  double **t_bufs[] = {&A->ntr, &A->tlo, &A->dQ, &A->aw_tr, &A->tb_tr, &A->xb_tr[0],
                       &A->xb_tr[1]};
  double **et_bufs[] = {&A->dflux, &A->r1, &A->r2};
  double **v_bufs[] = {&A->pv_vertex, &A->dvort, &A->dpvv, &A->vT, &A->hvT};
  for (auto b : e_bufs)
    if ((rc = dalloc(*A, E, b))) return rc;
  for (auto b : c_bufs)
    if ((rc = dalloc(*A, C, b))) return rc;
  for (auto b : t_bufs)
    if ((rc = dalloc(*A, T, b))) return rc;
  for (auto b : et_bufs)
    if ((rc = dalloc(*A, ET, b))) return rc;
  // hbf: one part of E x L x T per stencil slot (the kernels address each
  // part from its own base)
  if (M.nT) {
    if ((rc = dalloc_parts(*A, ET, std::max(1, M.wAdv), &A->hbf))) return rc;
    A->T.hslot = ET;
  }
  for (auto b : v_bufs)
    if ((rc = dalloc(*A, V, b))) return rc;
  for (int s = 0; s < 5; s++)
    if ((rc = dalloc(*A, E, &A->du[s])) || (rc = dalloc(*A, C, &A->dh[s])) ||
        (rc = dalloc(*A, T, &A->dT[s])))
      return rc;
  return 0;
}

int copy_d2d(mpas_ctx *ctx, double *dst, const double *src, size_t n,
             hipStream_t st = nullptr) {
  if (!n) return 0;
  MPAS_HIP(hipMemcpyAsync(dst, src, sizeof(double) * n, hipMemcpyDeviceToDevice,
                          st ? st : ctx->stream));
  return 0;
}

// The tape a replay writes: its records and its own replay buffers.
struct TapeView {
  std::vector<Rec> *rec;
  double *pv_vertex, *nu, *nh, *ntr, *tlo;
};
TapeView primary(AdjState &A) { return {&A.rec, A.pv_vertex, A.nu, A.nh, A.ntr, A.tlo}; }
TapeView spare(AdjState &A) {
  auto &t = A.spare;
  return {&t.rec, t.pv_vertex, t.nu, t.nh, t.ntr, t.tlo};
}
void swap_tapes(AdjState &A) {
  std::swap(A.rec, A.spare.rec);
  std::swap(A.pv_vertex, A.spare.pv_vertex);
  std::swap(A.nu, A.spare.nu);
  std::swap(A.nh, A.spare.nh);
  std::swap(A.ntr, A.spare.ntr);
  std::swap(A.tlo, A.spare.tlo);
}

// State (u, h, tr) into record r and the diagnostics of it, on stream st
// (nullptr: the context's stream).
int restore(mpas_ctx *ctx, const TapeView &tv, Rec &r, const double *u, const double *h,
            const double *tr, hipStream_t st = nullptr) {
  const DMesh &M = ctx->dm;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL;
  int rc;
  if ((rc = copy_d2d(ctx, r.u, u, E, st)) || (rc = copy_d2d(ctx, r.h, h, C, st)) ||
      (rc = copy_d2d(ctx, r.tr, tr, C * M.nT, st)))
    return rc;
  return diagnostics_dm(M, st ? st : ctx->stream, ctx->params.visc2 != 0.0, r.u, r.h, r.ke,
                        r.div, r.vort, tv.pv_vertex, r.he, r.pve);
}
int restore(mpas_ctx *ctx, AdjState &A, Rec &r, const double *u, const double *h,
            const double *tr) {
  return restore(ctx, primary(A), r, u, h, tr);
}

// One forward step from record 4i into records 4i+1 .. 4i+4 of the tape,
// on stream st (nullptr: the context's stream, with profiling).
int record_step(mpas_ctx *ctx, const TapeView &tv, int i, double dt, hipStream_t st = nullptr) {
  std::vector<Rec> &rec = *tv.rec;
  const Rec &r0 = rec[4 * i];
  for (int s = 0; s < 4; s++) {
    const Rec &x = rec[4 * i + s], &y = rec[4 * i + s + 1];
    StageIO S = stage_params(ctx, s, dt);
    S.xu = x.u;
    S.xh = x.h;
    S.xtr = x.tr;
    S.h_edge = x.he;
    S.ke = x.ke;
    S.pv_edge = x.pve;
    S.vort = x.vort;
    S.div = x.div;
    S.h_s = ctx->sw->h_s;
    S.u0 = r0.u;
    S.h0 = r0.h;
    S.tr0 = r0.tr;
    S.nu = tv.nu;
    S.nh = tv.nh;
    S.ntr = tv.ntr;
    S.yu = y.u;
    S.yh = y.h;
    S.ytr = y.tr;
    S.ke_y = y.ke;
    S.div_y = y.div;
    S.anti = x.anti;
    S.rfac = x.rfac;
    S.tlo = tv.tlo;
    const int rc = st ? launch_stage_dm(ctx->dm, st, S, y.vort, tv.pv_vertex, y.he, y.pve)
                      : launch_stage(ctx, S, y.vort, tv.pv_vertex, y.he, y.pve);
    if (rc) return rc;
  }
  return 0;
}
int record_step(mpas_ctx *ctx, AdjState &A, int i, double dt) {
  return record_step(ctx, primary(A), i, dt);
}

// Replays checkpoint segment m (len steps) into a tape.
int replay(mpas_ctx *ctx, AdjState &A, const TapeView &tv, int m, int len,
           hipStream_t st = nullptr) {
  if (int rc = restore(ctx, tv, (*tv.rec)[0], A.ck[3 * m], A.ck[3 * m + 1], A.ck[3 * m + 2], st))
    return rc;
  for (int i = 0; i < len; i++)
    if (int rc = record_step(ctx, tv, i, A.dt, st)) return rc;
  return 0;
}

// The spare tape, its stream and events, when the device has room for them
// (with 8 GiB to spare); returns 0 whether or not it was allocated.
int try_spare(mpas_ctx *ctx, AdjState &A) {
  if (A.spare_state) return 0;
  A.spare_state = -1;
  const DMesh &M = ctx->dm;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, V = size_t(M.nV) * M.nL;
  const size_t T = C * M.nT;
  const size_t per_rec = 3 * E + 3 * C + V + T + 2 * T + E * M.nT;
  const size_t need = 8 * (per_rec * A.rec.size() + V + E + C + 2 * T);
  size_t free_b = 0, total_b = 0;
  MPAS_HIP(hipMemGetInfo(&free_b, &total_b));
  if (free_b < need + (size_t(8) << 30)) return 0;
  A.spare.rec.resize(A.rec.size());
  int rc;
  for (Rec &r : A.spare.rec)
    if ((rc = alloc_rec(A, M, r))) return rc;
  if ((rc = dalloc(A, V, &A.spare.pv_vertex)) || (rc = dalloc(A, E, &A.spare.nu)) ||
      (rc = dalloc(A, C, &A.spare.nh)) || (rc = dalloc(A, T, &A.spare.ntr)) ||
      (rc = dalloc(A, T, &A.spare.tlo)))
    return rc;
  MPAS_HIP(hipStreamCreateWithFlags(&A.replay_stream, hipStreamNonBlocking));
  MPAS_HIP(hipEventCreateWithFlags(&A.ev_swept, hipEventDisableTiming));
  MPAS_HIP(hipEventCreateWithFlags(&A.ev_replayed, hipEventDisableTiming));
  A.spare_state = 1;
  return 0;
}

AdjParams adj_params(const mpas_ctx *ctx, int s, double dt) {
  const DMesh &M = ctx->dm;
  const mpas_params &Pm = ctx->params;
  AdjParams P{};
  P.s = s;
  P.nT = M.nT;
  P.freeze_u = Pm.freeze_u;
  P.use_visc = Pm.visc2 != 0.0;
  P.use_hmix = M.nT && Pm.kappa2 != 0.0;
  P.fct = M.nT && Pm.tracer_scheme == MPAS_TRACER_FCT;
  P.w_rk = rk_weight(s, dt);
  P.w_sub = rk_substep_weight(s, dt);
  P.w_rk_m = s > 0 ? rk_weight(s - 1, dt) : 0.0;
  P.w_sub_m = s > 0 ? rk_substep_weight(s - 1, dt) : 0.0;
  P.visc_coef = M.msd * Pm.visc2;
  P.kappa = Pm.kappa2;
  return P;
}

#define ADJ_LAUNCH_BS(bs, kernel, n, ...)                                                  \
  do {                                                                                     \
    if (int rc_ = prof_begin(ctx, #kernel, 0.0)) return rc_;                               \
    hipLaunchKernelGGL(kernel, dim3(xcd_grid(int64_t(n) * M.nL, bs)), dim3(bs), 0,         \
                       ctx->stream,                                                        \
                       __VA_ARGS__);                                                       \
    if (int rc_ = after_launch(#kernel)) return rc_;                                       \
    if (int rc_ = prof_end(ctx)) return rc_;                                               \
  } while (0)
#define ADJ_LAUNCH(kernel, n, ...) ADJ_LAUNCH_BS(A.ad_block, kernel, n, __VA_ARGS__)
#define TLM_LAUNCH(kernel, n, ...) ADJ_LAUNCH_BS(A.tl_block, kernel, n, __VA_ARGS__)

// TLM of one step along records 0..4: dX[0] -> dX[4].
int tlm_step(mpas_ctx *ctx, AdjState &A, double dt) {
  const DMesh &M = ctx->dm;
  for (int s = 0; s < 4; s++) {
    const AdjParams P = adj_params(ctx, s, dt);
    const Rec &R = A.rec[s], &Ry = A.rec[s + 1], &R0 = A.rec[0];
    TLM_LAUNCH(k_tl_vertex, M.nV, M, P, R, A.du[s], A.dh[s], A.dvort, A.dpvv);
    TLM_LAUNCH(k_tl_edge_a, M.nE, M, P, R, A.du[s], A.dh[s], A.dT[s], A.dpvv, A.dhe, A.dpve,
               A.dflux);
    TLM_LAUNCH(k_tl_cell, M.nC, M, P, R, Ry, R0, A.du[s], A.dh[0], A.dT[0], A.dhe, A.dflux,
               A.dnh, A.dQ, A.dke, A.ddiv, A.dh[s + 1], A.dT[s + 1]);
    TLM_LAUNCH(k_tl_edge_b, M.nE, M, P, R, A.du[s], A.du[0], A.dh[s], A.dhe, A.dpve, A.dke,
               A.ddiv, A.dvort, A.dnu, A.du[s + 1]);
  }
  return 0;
}

// Adjoint of step i of the tape (records 4i .. 4i+4): the adjoint of its
// final state in xb[src] -> the adjoint of its initial state in xb[1-src].
// prep: the step starts with k_ad_prep (else the previous step in the
// sweep, step i + 1, ran its prep); fuse_next: the stage-0 kernels run
// the prep of step i - 1.
int adj_step(mpas_ctx *ctx, AdjState &A, int i, double dt, int src, bool prep = true,
             bool fuse_next = false) {
  const DMesh &M = ctx->dm;
  const int dst = 1 - src;
  const Rec &R0 = A.rec[4 * i], &Rf = A.rec[4 * i + 4];
  if (prep) {
    const AdjParams P = adj_params(ctx, 3, dt);
    ADJ_LAUNCH(k_ad_prep, std::max(M.nE, M.nC), M, P, Rf, R0, A.xb_u[src], A.xb_h[src],
               A.xb_tr[src],
               PrepOut{A.xb_u[dst], A.xb_h[dst], A.xb_tr[dst], A.aw_u, A.aw_h, A.aw_tr, A.tb_u,
                       A.tb_h, A.tb_tr});
  }
  // step i - 1 sweeps from xb[dst] into xb[src]
  PrepNext N{};
  if (fuse_next && i > 0) {
    N.on = 1;
    N.R0n = A.rec[4 * (i - 1)];
    N.w_rk = adj_params(ctx, 3, dt).w_rk;
    N.Q = PrepOut{A.xb_u[src], A.xb_h[src], A.xb_tr[src], A.aw_u, A.aw_h, A.aw_tr, A.tb_u,
                  A.tb_h, A.tb_tr};
  }
  for (int s = 3; s >= 0; s--) {
    const AdjParams P = adj_params(ctx, s, dt);
    const Rec &R = A.rec[4 * i + s];
    // levels on lanes (kOneLaneLevels or more): one lane per entity and
    // level, each holding all of the sums' partials
    // levels on lanes with a multiple of 64 levels: one entity per wave
    const bool uni = M.nL >= kOneLaneLevels && M.nL % 64 == 0 && A.uni_ok;
    if (A.edge_fast) {
      if (uni)
        ADJ_LAUNCH((k_ad_edge_a<1, true, true>), M.nE, M, A.T, P, R, A.tb_u, A.tb_h, A.tb_tr,
                   A.ubA, A.heb, A.pveb, A.r1, A.r2, A.hbf);
      else if (M.nL >= kOneLaneLevels)
        ADJ_LAUNCH((k_ad_edge_a<1, true>), M.nE, M, A.T, P, R, A.tb_u, A.tb_h, A.tb_tr, A.ubA,
                   A.heb, A.pveb, A.r1, A.r2, A.hbf);
      else
        ADJ_LAUNCH((k_ad_edge_a<kEdgeLanes, true>), int64_t(M.nE) * kEdgeLanes, M, A.T, P, R,
                   A.tb_u, A.tb_h, A.tb_tr, A.ubA, A.heb, A.pveb, A.r1, A.r2, A.hbf);
    } else if (M.nL >= kOneLaneLevels) {
      ADJ_LAUNCH((k_ad_edge_a<1, false>), M.nE, M, A.T, P, R, A.tb_u, A.tb_h, A.tb_tr, A.ubA,
                 A.heb, A.pveb, A.r1, A.r2, A.hbf);
    } else {
      ADJ_LAUNCH((k_ad_edge_a<kEdgeLanes, false>), int64_t(M.nE) * kEdgeLanes, M, A.T, P, R,
                 A.tb_u, A.tb_h, A.tb_tr, A.ubA, A.heb, A.pveb, A.r1, A.r2, A.hbf);
    }
    const PrepNext Ns = s == 0 ? N : PrepNext{};
    const CellArgs G{A.tb_u, A.heb,    A.hvT,      A.pveb,       A.r1,   A.r2,   A.hbf, A.keT,
                     A.divT, A.xb_h[dst], A.xb_tr[dst], A.aw_h, A.aw_tr, A.tb_h, A.tb_tr};
    if (M.nL >= kOneLaneLevels) {
      ADJ_LAUNCH(k_ad_vertex, M.nV, M, P, R, A.pveb, A.tb_u, A.vT, A.hvT);
      if (uni)
        ADJ_LAUNCH((k_ad_cell<1, true>), M.nC, M, A.T, P, R, R0, G, Ns);
      else
        ADJ_LAUNCH(k_ad_cell<1>, M.nC, M, A.T, P, R, R0, G, Ns);
    } else if (A.merge_vertex) {
      // one launch: vertex workgroups first, then the cells'
      const int bs = A.ad_block;
      const int vb = xcd_grid(int64_t(M.nV) * M.nL, bs);
      const int cb = xcd_grid(int64_t(M.nC) * kCellLanes * M.nL, bs);
      if (int rc_ = prof_begin(ctx, "k_ad_cell_vertex", 0.0)) return rc_;
      hipLaunchKernelGGL(k_ad_cell_vertex<kCellLanes>, dim3(vb + cb), dim3(bs), 0, ctx->stream,
                         M, A.T, P, R, R0, G, Ns, A.vT, vb);
      if (int rc_ = after_launch("k_ad_cell_vertex")) return rc_;
      if (int rc_ = prof_end(ctx)) return rc_;
    } else {
      ADJ_LAUNCH(k_ad_vertex, M.nV, M, P, R, A.pveb, A.tb_u, A.vT, A.hvT);
      ADJ_LAUNCH(k_ad_cell<kCellLanes>, int64_t(M.nC) * kCellLanes, M, A.T, P, R, R0, G, Ns);
    }
    ADJ_LAUNCH(k_ad_edge_b, M.nE, M, P, R, A.ubA, A.keT, A.divT, A.vT, A.xb_u[dst], A.aw_u,
               A.tb_u, Ns);
  }
  return 0;
}

// The reverse sweep of the len steps the tape holds, from the adjoint in
// xb[src]: launched directly, or replayed from a hipGraph captured the first
// time for this tape, parity, length and time step.
int sweep_segment(mpas_ctx *ctx, AdjState &A, int len, int src) {
  // each step but the segment's last in the sweep runs the next step's
  // prep in its stage-0 kernels
  auto direct = [&]() -> int {
    int x = src;
    for (int i = len - 1; i >= 0; i--) {
      if (int rc = adj_step(ctx, A, i, A.dt, x, i == len - 1 || !A.fuse_prep, A.fuse_prep))
        return rc;
      x = 1 - x;
    }
    return 0;
  };
  if (!A.graph || !A.graph_ok || ctx->prof.on) return direct();
  // the tape's buffers: only records 0 and 4 ever trade places (the TLM's
  // step), so these two name the whole set
  const auto key = std::make_tuple(static_cast<const void *>(A.rec[0].u),
                                   static_cast<const void *>(A.rec[4].u), src, len);
  auto it = A.sweeps.find(key);
  if (it != A.sweeps.end()) {
    MPAS_HIP(hipGraphLaunch(it->second, ctx->stream));
    return 0;
  }
  MPAS_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  const int rc = direct();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess)
    MPAS_FAIL(MPAS_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess)
    MPAS_FAIL(MPAS_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  A.sweeps[key] = x;
  MPAS_HIP(hipGraphLaunch(x, ctx->stream));
  return 0;
}

int check_adj(mpas_ctx *ctx) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->adj) MPAS_FAIL(MPAS_ESTATE, "mpas_adj_init has not been called");
  if (!ctx->sw || !ctx->sw->ready)
    MPAS_FAIL(MPAS_ESTATE, "upload a state first (mpas_sw_state_upload sets h_s)");
  return 0;
}

int check_state_ptrs(const mpas_ctx *ctx, const void *u, const void *h, const void *tr,
                     const char *what) {
  if (!u || !h || (ctx->dm.nT && !tr))
    MPAS_FAIL(MPAS_EINVAL, std::string(what) + ": NULL state pointer");
  return 0;
}

}  // namespace
}  // namespace mpasx

using namespace mpasx;

extern "C" {

int mpas_adj_init(mpas_ctx *ctx, int checkpoint_every) {
  if (int rc = check_ctx(ctx)) return rc;
  if (checkpoint_every < 1) MPAS_FAIL(MPAS_EINVAL, "mpas_adj_init: checkpoint_every must be >= 1");
  const DMesh &M = ctx->dm;
  const mpas_params &P = ctx->params;
  if (M.nT && P.tracer_scheme != MPAS_TRACER_FCT)
    MPAS_FAIL(MPAS_EINVAL, "the adjoint covers the FCT tracer scheme (or no tracers)");
  if (M.nT > 8) MPAS_FAIL(MPAS_ESHAPE, "the adjoint supports at most 8 tracers");
  if (M.nT && (M.mlc || M.hoMask || !ctx->aligned_c2c))
    MPAS_FAIL(MPAS_EINVAL, "the adjoint needs every level active, a full high-order mask and "
                           "cellsOnCell across edgesOnCell");
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  adj_free(ctx);
  if (int rc = adj_alloc(ctx, checkpoint_every)) {
    adj_free(ctx);
    return rc;
  }
  MPAS_HIP(hipDeviceSynchronize());
  return 0;
}

int mpas_adj_forward(mpas_ctx *ctx, double dt, int nsteps, const double *u0, const double *h0,
                     const double *tr0, double *u1, double *h1, double *tr1) {
  if (int rc = check_adj(ctx)) return rc;
  if (int rc = check_state_ptrs(ctx, u0, h0, tr0, "mpas_adj_forward")) return rc;
  if (nsteps < 1) MPAS_FAIL(MPAS_EINVAL, "mpas_adj_forward: nsteps must be >= 1");
  AdjState &A = *ctx->adj;
  SwState &W = *ctx->sw;
  const DMesh &M = ctx->dm;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, T = C * M.nT;
  const int nseg = (nsteps + A.K - 1) / A.K;
  while (int(A.ck.size()) < 3 * nseg) {
    double *u, *h, *tr;
    int rc;
    if ((rc = dalloc(A, E, &u)) || (rc = dalloc(A, C, &h)) || (rc = dalloc(A, std::max<size_t>(T, 1), &tr)))
      return rc;
    A.ck.push_back(u);
    A.ck.push_back(h);
    A.ck.push_back(tr);
  }
  int rc;
  if ((rc = copy_d2d(ctx, W.u, u0, E)) || (rc = copy_d2d(ctx, W.h, h0, C)) ||
      (rc = copy_d2d(ctx, W.tr, tr0, T)))
    return rc;
  W.cur = 0;
  W.elapsed = 0;
  if ((rc = sw_diagnostics(ctx, W.u, W.h, W.cur))) return rc;
  for (int m = 0; m < nseg; m++) {
    const int len = std::min(A.K, nsteps - m * A.K);
    if ((rc = copy_d2d(ctx, A.ck[3 * m], W.u, E)) || (rc = copy_d2d(ctx, A.ck[3 * m + 1], W.h, C)) ||
        (rc = copy_d2d(ctx, A.ck[3 * m + 2], W.tr, T)))
      return rc;
    if (m < nseg - 1) {
      if ((rc = sw_rk4_steps(ctx, dt, len))) return rc;
      continue;
    }
    // the last segment runs into the tape (the backward sweep starts on it
    // without a replay); the same kernels, so the same state bit for bit
    if ((rc = restore(ctx, A, A.rec[0], W.u, W.h, W.tr))) return rc;
    for (int i = 0; i < len; i++)
      if ((rc = record_step(ctx, A, i, dt))) return rc;
    const Rec &f = A.rec[4 * len];
    if ((rc = copy_d2d(ctx, W.u, f.u, E)) || (rc = copy_d2d(ctx, W.h, f.h, C)) ||
        (rc = copy_d2d(ctx, W.tr, f.tr, T)) || (rc = sw_diagnostics(ctx, W.u, W.h, W.cur)))
      return rc;
    W.elapsed += dt * len;
  }
  if ((u1 && (rc = copy_d2d(ctx, u1, W.u, E))) || (h1 && (rc = copy_d2d(ctx, h1, W.h, C))) ||
      (tr1 && (rc = copy_d2d(ctx, tr1, W.tr, T))))
    return rc;
  A.nsteps = nsteps;
  if (A.dt != dt) drop_sweeps(A);  // the sweeps hold the stage weights
  A.dt = dt;
  A.forward_done = true;
  A.taped = 0;
  A.tape_seg = nseg - 1;
  return 0;
}

int mpas_adj_backward(mpas_ctx *ctx, const double *u1_bar, const double *h1_bar,
                      const double *tr1_bar, double *u0_bar, double *h0_bar, double *tr0_bar) {
  if (int rc = check_adj(ctx)) return rc;
  if (int rc = check_state_ptrs(ctx, u1_bar, h1_bar, tr1_bar, "mpas_adj_backward")) return rc;
  if (int rc = check_state_ptrs(ctx, u0_bar, h0_bar, tr0_bar, "mpas_adj_backward")) return rc;
  AdjState &A = *ctx->adj;
  if (!A.forward_done) MPAS_FAIL(MPAS_ESTATE, "mpas_adj_backward before mpas_adj_forward");
  const DMesh &M = ctx->dm;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, T = C * M.nT;
  int rc, src = 0;
  if ((rc = copy_d2d(ctx, A.xb_u[0], u1_bar, E)) || (rc = copy_d2d(ctx, A.xb_h[0], h1_bar, C)) ||
      (rc = copy_d2d(ctx, A.xb_tr[0], tr1_bar, T)))
    return rc;
  const int nseg = (A.nsteps + A.K - 1) / A.K;
  auto seg_len = [&](int m) { return std::min(A.K, A.nsteps - m * A.K); };
  if (nseg > 1 && A.overlap && (rc = try_spare(ctx, A))) return rc;
  const bool overlap = nseg > 1 && A.overlap && A.spare_state == 1;
  hipStream_t rs = A.replay_stream;
  if (overlap) {  // the replays read checkpoints the context's stream wrote
    MPAS_HIP(hipEventRecord(A.ev_swept, ctx->stream));
    MPAS_HIP(hipStreamWaitEvent(rs, A.ev_swept, 0));
  }
  if (A.tape_seg != nseg - 1) {
    if ((rc = replay(ctx, A, primary(A), nseg - 1, seg_len(nseg - 1)))) return rc;
    A.tape_seg = nseg - 1;
  }
  for (int m = nseg - 1; m >= 0; m--) {
    const int len = seg_len(m);
    if (m != A.tape_seg) {  // no spare tape: replay the segment into the tape
      if ((rc = replay(ctx, A, primary(A), m, len))) return rc;
      A.tape_seg = m;
    }
    // the next segment's replay runs beside this segment's sweep; the spare
    // tape's last reader (the previous sweep) finished before ev_swept
    const bool ahead = overlap && m > 0;
    if (ahead && (rc = replay(ctx, A, spare(A), m - 1, seg_len(m - 1), rs))) return rc;
    if (ahead) MPAS_HIP(hipEventRecord(A.ev_replayed, rs));
    if ((rc = sweep_segment(ctx, A, len, src))) return rc;
    if (len % 2) src = 1 - src;
    A.taped = len;
    if (ahead) {
      MPAS_HIP(hipEventRecord(A.ev_swept, ctx->stream));
      MPAS_HIP(hipStreamWaitEvent(rs, A.ev_swept, 0));
      MPAS_HIP(hipStreamWaitEvent(ctx->stream, A.ev_replayed, 0));
      swap_tapes(A);
      A.tape_seg = m - 1;
    }
  }
  if ((rc = copy_d2d(ctx, u0_bar, A.xb_u[src], E)) || (rc = copy_d2d(ctx, h0_bar, A.xb_h[src], C)) ||
      (rc = copy_d2d(ctx, tr0_bar, A.xb_tr[src], T)))
    return rc;
  return 0;
}

int mpas_tlm_forward(mpas_ctx *ctx, double dt, int nsteps, const double *u0, const double *h0,
                     const double *tr0, const double *du0, const double *dh0, const double *dtr0,
                     double *u1, double *h1, double *tr1, double *du1, double *dh1, double *dtr1) {
  if (int rc = check_adj(ctx)) return rc;
  if (int rc = check_state_ptrs(ctx, u0, h0, tr0, "mpas_tlm_forward")) return rc;
  if (int rc = check_state_ptrs(ctx, du0, dh0, dtr0, "mpas_tlm_forward")) return rc;
  if (nsteps < 1) MPAS_FAIL(MPAS_EINVAL, "mpas_tlm_forward: nsteps must be >= 1");
  AdjState &A = *ctx->adj;
  const DMesh &M = ctx->dm;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, T = C * M.nT;
  int rc;
  A.taped = 0;
  A.tape_seg = -1;
  if ((rc = restore(ctx, A, A.rec[0], u0, h0, tr0))) return rc;
  if ((rc = copy_d2d(ctx, A.du[0], du0, E)) || (rc = copy_d2d(ctx, A.dh[0], dh0, C)) ||
      (rc = copy_d2d(ctx, A.dT[0], dtr0, T)))
    return rc;
  for (int step = 0; step < nsteps; step++) {
    if ((rc = record_step(ctx, A, 0, dt))) return rc;
    if ((rc = tlm_step(ctx, A, dt))) return rc;
    std::swap(A.rec[0], A.rec[4]);
    std::swap(A.du[0], A.du[4]);
    std::swap(A.dh[0], A.dh[4]);
    std::swap(A.dT[0], A.dT[4]);
  }
  const Rec &r = A.rec[0];
  if ((u1 && (rc = copy_d2d(ctx, u1, r.u, E))) || (h1 && (rc = copy_d2d(ctx, h1, r.h, C))) ||
      (tr1 && (rc = copy_d2d(ctx, tr1, r.tr, T))) || (du1 && (rc = copy_d2d(ctx, du1, A.du[0], E))) ||
      (dh1 && (rc = copy_d2d(ctx, dh1, A.dh[0], C))) ||
      (dtr1 && (rc = copy_d2d(ctx, dtr1, A.dT[0], T))))
    return rc;
  return 0;
}

int mpas_adj_frozen_forward(mpas_ctx *ctx, int nsteps, const double *u0, const double *h0,
                            const double *tr0, double *u1, double *h1, double *tr1) {
  if (int rc = check_adj(ctx)) return rc;
  if (int rc = check_state_ptrs(ctx, u0, h0, tr0, "mpas_adj_frozen_forward")) return rc;
  AdjState &A = *ctx->adj;
  if (nsteps < 1 || nsteps > A.taped)
    MPAS_FAIL(MPAS_ESTATE, "mpas_adj_frozen_forward: the tape holds " + std::to_string(A.taped) +
                               " steps (run mpas_adj_backward on at most checkpoint_every steps)");
  SwState &W = *ctx->sw;
  const DMesh &M = ctx->dm;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, T = C * M.nT;
  int rc;
  if ((rc = copy_d2d(ctx, W.u, u0, E)) || (rc = copy_d2d(ctx, W.h, h0, C)) ||
      (rc = copy_d2d(ctx, W.tr, tr0, T)))
    return rc;
  W.cur = 0;
  W.elapsed = 0;
  if ((rc = sw_diagnostics(ctx, W.u, W.h, W.cur))) return rc;
  std::vector<double *> fr(size_t(4) * nsteps);
  for (int i = 0; i < nsteps; i++)
    for (int s = 0; s < 4; s++) fr[4 * i + s] = A.rec[4 * i + s].rfac;
  if ((rc = sw_rk4_steps(ctx, A.dt, nsteps, fr.data()))) return rc;
  if ((u1 && (rc = copy_d2d(ctx, u1, W.u, E))) || (h1 && (rc = copy_d2d(ctx, h1, W.h, C))) ||
      (tr1 && (rc = copy_d2d(ctx, tr1, W.tr, T))))
    return rc;
  return 0;
}

int mpas_adj_config(mpas_ctx *ctx, int ad_block, int tl_block, int overlap, int graph) {
  if (int rc = check_ctx(ctx)) return rc;
  if (!ctx->adj) MPAS_FAIL(MPAS_ESTATE, "mpas_adj_init has not been called");
  for (int b : {ad_block, tl_block})
    if (b < 64 || b > kAdjMaxBlock || b % 64)
      MPAS_FAIL(MPAS_EINVAL, "mpas_adj_config: workgroup sizes are multiples of 64 up to 512");
  AdjState &A = *ctx->adj;
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  if (A.ad_block != ad_block) drop_sweeps(A);
  A.ad_block = ad_block;
  A.tl_block = tl_block;
  A.overlap = overlap != 0;
  A.graph = graph != 0 && A.graph_ok;
  return 0;
}

int mpas_adj_get_config(const mpas_ctx *ctx, int *ad_block, int *tl_block, int *overlap,
                        int *graph) {
  if (!ctx || !ctx->adj) MPAS_FAIL(MPAS_ESTATE, "mpas_adj_init has not been called");
  const AdjState &A = *ctx->adj;
  if (ad_block) *ad_block = A.ad_block;
  if (tl_block) *tl_block = A.tl_block;
  if (overlap) *overlap = A.overlap;
  if (graph) *graph = A.graph;
  return 0;
}

int mpas_adj_autotune(mpas_ctx *ctx, double dt, int nsteps, char *report, size_t cap) {
  if (int rc = check_adj(ctx)) return rc;
  if (nsteps < 1) MPAS_FAIL(MPAS_EINVAL, "mpas_adj_autotune: nsteps must be >= 1");
  if (!(dt > 0)) MPAS_FAIL(MPAS_EINVAL, "mpas_adj_autotune: dt must be positive");
  AdjState &A = *ctx->adj;
  SwState &W = *ctx->sw;
  const DMesh &M = ctx->dm;
  const size_t E = size_t(M.nE) * M.nL, C = size_t(M.nC) * M.nL, T = C * M.nT;
  MPAS_HIP(hipStreamSynchronize(ctx->stream));
  // the trials start from the resident state, which they advance: keep it
  // (x0) and restore it after; y and xb are the sweep's input and output
  std::vector<void *> tmp;
  auto get = [&](size_t n, double **p) -> int {
    *p = nullptr;
    if (hipMalloc(reinterpret_cast<void **>(p), sizeof(double) * std::max<size_t>(n, 1)) !=
        hipSuccess)
      MPAS_FAIL(MPAS_ENOMEM, "mpas_adj_autotune: no memory for the trial state");
    tmp.push_back(*p);
    return 0;
  };
  double *x0[3], *y[3], *xb[3], *dx[3];
  int rc = 0;
  const size_t n3[3] = {E, C, T};
  for (int k = 0; k < 3 && !rc; k++)
    if (!(rc = get(n3[k], &x0[k])) && !(rc = get(n3[k], &y[k])) && !(rc = get(n3[k], &xb[k])))
      rc = get(n3[k], &dx[k]);
  const double *cur[3] = {W.u, W.h, W.tr};
  for (int k = 0; k < 3 && !rc; k++)
    if (n3[k] && hipMemcpyAsync(x0[k], cur[k], sizeof(double) * n3[k], hipMemcpyDeviceToDevice,
                                ctx->stream) != hipSuccess)
      rc = MPAS_EHIP;
  std::string log = "[";
  // seconds of one gradient (forward with checkpoints + sweep) or one TLM run
  auto timed = [&](bool tlm) -> double {
    if (rc) return 1e30;
    double best = 1e30;
    // the best of the runs after the first, which warms up (graph capture)
    for (int rep = 0; rep < 3 && !rc; rep++) {
      (void)hipStreamSynchronize(ctx->stream);
      const auto t0 = std::chrono::steady_clock::now();
      if (tlm)
        rc = mpas_tlm_forward(ctx, dt, nsteps, x0[0], x0[1], x0[2], x0[0], x0[1], x0[2],
                              nullptr, nullptr, nullptr, dx[0], dx[1], dx[2]);
      else if (!(rc = mpas_adj_forward(ctx, dt, nsteps, x0[0], x0[1], x0[2], y[0], y[1], y[2])))
        rc = mpas_adj_backward(ctx, y[0], y[1], y[2], xb[0], xb[1], xb[2]);
      if (hipStreamSynchronize(ctx->stream) != hipSuccess && !rc) rc = MPAS_EHIP;
      if (rep > 0)
        best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                                  .count());
    }
    return rc ? 1e30 : best;
  };
  auto note = [&](const char *what, double t) {
    char line[200];
    std::snprintf(line, sizeof line,
                  "%s{\"trial\": \"%s\", \"ad_block\": %d, \"tl_block\": %d, "
                  "\"overlap\": %d, \"graph\": %d, \"ms\": %.4f}",
                  log.size() > 1 ? ", " : "", what, A.ad_block, A.tl_block, A.overlap, A.graph,
                  1e3 * t);
    log += line;
  };
  const int sizes[] = {64, 128, 256, 512};
  double best = timed(false);
  note("start", best);
  // the schedule: replays beside the sweep or in turn, sweeps from graphs or not
  for (int ov = 0; ov < 2; ov++)
    for (int gr = 0; gr < 1 + A.graph_ok; gr++) {
      if (ov == A.overlap && gr == A.graph) continue;
      const int po = A.overlap, pg = A.graph;
      A.overlap = ov;
      A.graph = gr;
      const double t = timed(false);
      note("schedule", t);
      if (t < best) best = t;
      else A.overlap = po, A.graph = pg;
    }
  // the tile of the adjoint kernels
  for (int bs : sizes) {
    if (bs == A.ad_block) continue;
    const int keep = A.ad_block;
    drop_sweeps(A);
    A.ad_block = bs;
    const double t = timed(false);
    note("adjoint block", t);
    if (t < best) best = t;
    else drop_sweeps(A), A.ad_block = keep;
  }
  // the tile of the tangent-linear kernels, on TLM runs
  double tbest = timed(true);
  note("tlm start", tbest);
  for (int bs : sizes) {
    if (bs == A.tl_block) continue;
    const int keep = A.tl_block;
    A.tl_block = bs;
    const double t = timed(true);
    note("tlm block", t);
    if (t < tbest) tbest = t;
    else A.tl_block = keep;
  }
  // the resident state as it was; the tape holds the trials' runs
  for (int k = 0; k < 3 && !rc; k++)
    if (n3[k] && hipMemcpyAsync(const_cast<double *>(cur[k]), x0[k], sizeof(double) * n3[k],
                                hipMemcpyDeviceToDevice, ctx->stream) != hipSuccess)
      rc = MPAS_EHIP;
  if (!rc) rc = sw_diagnostics(ctx, W.u, W.h, W.cur);
  (void)hipStreamSynchronize(ctx->stream);
  for (void *p : tmp) (void)hipFree(p);
  A.forward_done = false;
  A.taped = 0;
  A.tape_seg = -1;
  if (rc) return rc;
  log += "]";
  if (report && cap) {
    if (log.size() + 1 > cap) MPAS_FAIL(MPAS_ESHAPE, "mpas_adj_autotune: report buffer too small");
    std::memcpy(report, log.c_str(), log.size() + 1);
  }
  return 0;
}

}  // extern "C"
