// Ensembles: many members, on meshes of uneven size, advanced together.
// (The reference runs one mesh per process; this has no counterpart there.)
//
// Members that share a mesh and a time step form a group, stored
// member-minor: member g of a group is level g of one multi-level state
// ([entity][member], tracers [cell][member][tracer]).  Every kernel of the
// path treats levels independently, so a member evolves bit for bit as a
// single run of it would.  The kernels run with levels on lanes
// (stage_kernels.h): neighbouring lanes take neighbouring members of one
// entity, so state accesses and neighbour gathers are contiguous runs
// across members and each stencil load is shared by the members.
//
// Every phase of an RK stage (momentum, high-order flux, cells, tracer
// update, vertex and edge diagnostics) is ONE launch over all groups: a
// table maps each logical workgroup to its group and block, and the
// group's mesh view and stage descriptors are read from device memory.  A
// step is 24 launches, captured once in a hipGraph and replayed.
//
// Tunables (mpas_batch_autotune picks them by timing graph replays): the
// workgroup size of each of the 6 kernels, members per lane, the order of
// the groups' workgroups in the launch space, and XCD-contiguous placement.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_cooperative_groups.h>

#include "ctx.h"
#include "stage_kernels.h"

namespace mpasx {

// Device-side descriptor of one group.
struct BatchGroup {
  DMesh M;  // the mesh context's view with nL = members of the group
  StageIO S[4];  // stage s of a step (the ping-pong roles repeat every step)
  double *vort_y[4], *h_edge_y[4], *pv_edge_y[4];
  double *pv_vertex;
};

// Buffers of one group (the SwState fields, nL = members).
struct GroupBufs {
  double *u, *h, *tr, *nu, *nh, *ntr, *pu[2], *ph[2], *ptr[2];
  double *h_edge[2], *ke[2], *pv_edge[2], *vort[2], *div[2], *pv_vertex, *h_s;
  double *anti, *rfac, *tlo;
};

struct GroupInfo {
  int ctx = 0;   // mesh context
  double dt = 0;
  std::vector<int> members;
  std::vector<double> h_s;  // bottom topography shared by the group (set by the first upload)
  GroupBufs b{};
};

constexpr int kBatchKernels = 6;  // edges, ho, cells, tracer, vertex, edge_diag
constexpr int kBatchModes = 6;    // layouts of a group (see MODE below)
constexpr int kBatchMaxBlock = 512;

struct BatchArgs {
  const BatchGroup *__restrict__ grp;
  const int2 *__restrict__ tab;  // logical workgroup -> (group, its block); group -1: padding
  int xcd;                       // XCD-contiguous placement
};

}  // namespace mpasx

struct mpas_batch {
  std::vector<mpas_ctx *> meshes;
  int nmem = 0, nT = 0, narrow = 1, device = 0;
  std::vector<int> grp_of, idx_in_grp;  // per member
  std::vector<mpasx::GroupInfo> groups;
  std::vector<mpasx::BatchGroup> desc;
  std::vector<int> ready;
  hipStream_t stream = nullptr;
  void *arena = nullptr;
  size_t arena_bytes = 0;
  double *staging = nullptr;  // one member's largest field (uploads, downloads)
  mpasx::BatchGroup *d_grp = nullptr;
  // launch configuration and its tables
  int block[mpasx::kBatchKernels] = {256, 256, 256, 256, 256, 256};
  // per kernel and layout, set by the autotuner only (0: block[k])
  int mblock[mpasx::kBatchKernels][mpasx::kBatchModes] = {};
  int lpl = 1;    // members per lane
  // one entity per wave (members padded to whole waves): 0 for no group,
  // 1 for the groups that keep at least 3/4 of the padded lanes busy, 2 for
  // every group
  int wave_entity = 1;
  int order = 0;  // 0: group after group, 1: the groups' workgroups interleaved
  int xcd = 1;
  // workgroup tables per kernel and layout (index mode - 1); grid 0: no
  // group of that layout
  int2 *d_tab[mpasx::kBatchKernels][mpasx::kBatchModes] = {};
  int grid[mpasx::kBatchKernels][mpasx::kBatchModes] = {};
  // the persistent launch's tables: per kernel, (layout, group, block) of
  // blocks of 256 items over every layout
  int4 *d_ptab[mpasx::kBatchKernels] = {};
  int pn[mpasx::kBatchKernels] = {};
  int cus = 0;          // compute units of the device
  int tuned_mode = 0;   // the launch mode the autotuner found fastest
  // mesh order of the groups (MPAS_ORDER_*): per order and mesh, the context
  // the groups run on (order 0: the caller's; others owned, created on first
  // use) and, on the device, the original id of each cell and edge
  int mesh_order = 0;
  std::vector<mpas_ctx *> octx[3];
  std::vector<int32_t *> d_cell_of[3], d_edge_of[3];
  bool tables_ok = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace mpasx {
namespace {

__device__ __forceinline__ int batch_block(const BatchArgs &A) {
  const int b = blockIdx.x;
  if (!A.xcd) return b;
  const int per = gridDim.x / kXcds;
  return (b % kXcds) * per + b / kXcds;
}

#define BATCH_PROLOGUE                                                    \
  const int2 t = A.tab[batch_block(A)];                                   \
  if (t.x < 0) return;                                                    \
  const BatchGroup &g = A.grp[t.x];                                       \
  const int item = t.y * int(blockDim.x) + int(threadIdx.x)

// MODE: the groups' layout, each with its own launch:
//   1  levels on lanes, one member per lane;
//   2  the same with one entity per wave;
//   3, 4  1 and 2 with two members per lane, for the kernels in
//         level-group form (an even member count);
//   5, 6  1 and 2 with members_per_lane > 1 (the last chunk may be short).
// Layouts 1-4 fix the members per lane at compile time (one pass of the
// level loop, fewer registers); 5 and 6 read it from the descriptor.  The
// group's descriptors are copied to locals first: read through the pointer
// they would be reloaded after every store the bodies make (the stores
// could alias them).
constexpr int lol_of(int mode) { return mode % 2 == 0 ? 2 : 1; }
constexpr int lp_of(int mode) { return mode == 3 || mode == 4 ? 2 : 1; }
constexpr int fix_of(int mode) { return mode <= 4 ? lp_of(mode) : 0; }
// The stage descriptor of layout MODE: level pairs halve the chunks.
template <int MODE>
__device__ __forceinline__ StageIO stage_of(const BatchGroup &g, int s) {
  StageIO S = g.S[s];
  if (lp_of(MODE) == 2) {
    S.lpl = 2;
    S.nchunk = S.nchunk / 2;
  }
  return S;
}
// The momentum kernel gathers its edgesOnEdge slots in phases of 4 (the
// one-context kernels take 5): with two members per lane it then needs 157
// VGPRs instead of 171, within the 3-wave budget (ensemble 12.24 -> 12.18
// ms/step); the 3-wave hint stays as a guard.
#ifndef MPAS_BATCH_EDGES_WPE
#define MPAS_BATCH_EDGES_WPE 3
#endif
#if MPAS_BATCH_EDGES_WPE > 0
#define MPAS_BATCH_EDGES_WAVES __attribute__((amdgpu_waves_per_eu(MPAS_BATCH_EDGES_WPE, 8)))
#else
#define MPAS_BATCH_EDGES_WAVES
#endif
#ifndef MPAS_BATCH_QCH
#define MPAS_BATCH_QCH 4
#endif
// Advection slots per scheduling phase of the high-order flux kernels
// (0: all at once).
#ifndef MPAS_BATCH_QHO
#define MPAS_BATCH_QHO 5
#endif
// The phase bodies of layout MODE for one work item of group g, stage s.
template <int WE, int MODE>
__device__ __forceinline__ void run_edges(const BatchGroup &g, int s, int item) {
  const DMesh M = g.M;
  const StageIO S = stage_of<MODE>(g, s);
  stage_edges<WE, lol_of(MODE), lp_of(MODE), fix_of(MODE), MPAS_BATCH_QCH>(M, S, item);
}
template <int NT, int WA, int MODE>
__device__ __forceinline__ void run_ho(const BatchGroup &g, int s, int item) {
  const DMesh M = g.M;
  const StageIO S = stage_of<MODE>(g, s);
  stage_ho<NT, WA, lol_of(MODE), 0, lp_of(MODE), fix_of(MODE), false, MPAS_BATCH_QHO>(M, S, item);
}
template <int NT, int W, int MODE, int SCH>
__device__ __forceinline__ void run_cells(const BatchGroup &g, int s, int item) {
  if constexpr (lp_of(MODE) == 1) {  // never run with pairs
    const DMesh M = g.M;
    const StageIO S = g.S[s];
    stage_cells<NT, W, lol_of(MODE), 0, SCH, fix_of(MODE)>(M, S, item);
  }
}
template <int NT, int W, int MODE>
__device__ __forceinline__ void run_tracer(const BatchGroup &g, int s, int item) {
  if constexpr (lp_of(MODE) == 1) {  // never run with pairs
    const DMesh M = g.M;
    const StageIO S = g.S[s];
    stage_tracer<NT, W, lol_of(MODE), 0, fix_of(MODE)>(M, S, item);
  }
}

template <int WE, int MODE>
__global__ __launch_bounds__(kBatchMaxBlock) MPAS_BATCH_EDGES_WAVES void
kb_stage_edges(BatchArgs A, int s) {
  BATCH_PROLOGUE;
  run_edges<WE, MODE>(g, s, item);
}
template <int NT, int WA, int MODE>
__global__ __launch_bounds__(kBatchMaxBlock) void kb_stage_ho(BatchArgs A, int s) {
  BATCH_PROLOGUE;
  run_ho<NT, WA, MODE>(g, s, item);
}
template <int NT, int W, int MODE, int SCH>
__global__ __launch_bounds__(kBatchMaxBlock) void kb_stage_cells(BatchArgs A, int s) {
  if constexpr (lp_of(MODE) == 1) {  // never launched with pairs
    BATCH_PROLOGUE;
    run_cells<NT, W, MODE, SCH>(g, s, item);
  }
}
template <int NT, int W, int MODE>
__global__ __launch_bounds__(kBatchMaxBlock) void kb_stage_tracer(BatchArgs A, int s) {
  if constexpr (lp_of(MODE) == 1) {  // never launched with pairs
    BATCH_PROLOGUE;
    run_tracer<NT, W, MODE>(g, s, item);
  }
}
template <int MODE>
__device__ __forceinline__ void diag_vertex_lanes(const BatchGroup &g, int s, int item) {
  int sub;
  uint32_t k0, k1;
  const DMesh M = g.M;
  const StageIO S = stage_of<MODE>(g, s);
  double *const vort = g.vort_y[s], *const pvv = g.pv_vertex;
  const int v = split_levels<lol_of(MODE), 1, fix_of(MODE)>(S, M.nL, item, sub, k0, k1);
  if (v < M.nV) diag_vertex<lp_of(MODE)>(M, v, S.yh, S.yu, nullptr, nullptr, vort, pvv, k0, k1);
}
template <int MODE>
__device__ __forceinline__ void diag_edge_lanes(const BatchGroup &g, int s, int item) {
  int sub;
  uint32_t k0, k1;
  const DMesh M = g.M;
  const StageIO S = stage_of<MODE>(g, s);
  const double *const pvv = g.pv_vertex;
  double *const he = g.h_edge_y[s], *const pve = g.pv_edge_y[s];
  const int e = split_levels<lol_of(MODE), 1, fix_of(MODE)>(S, M.nL, item, sub, k0, k1);
  if (e < M.nE) diag_edge<lp_of(MODE)>(M, e, S.yh, pvv, he, pve, k0, k1);
}
template <int MODE>
__global__ __launch_bounds__(kBatchMaxBlock) void kb_diag_vertex(BatchArgs A, int s) {
  BATCH_PROLOGUE;
  diag_vertex_lanes<MODE>(g, s, item);
}
template <int MODE>
__global__ __launch_bounds__(kBatchMaxBlock) void kb_diag_edge(BatchArgs A, int s) {
  BATCH_PROLOGUE;
  diag_edge_lanes<MODE>(g, s, item);
}
#undef BATCH_PROLOGUE

// ---- one persistent launch ----------------------------------------------------
// Every phase of every stage of nsteps steps in one cooperative launch: the
// workgroups walk phase k's work table (layout, group, block of kPBlock
// items) block-stride, then wait at a grid barrier for the phase's results.
// The bodies are those of the phase kernels above, so every member evolves
// exactly as with the per-phase launches.  Instantiated for the shapes of
// the ensemble workloads (meshes of at most 6 edges per cell, 0, 1, 2 or 4
// tracers, FCT).
constexpr int kPBlock = 256;
#if MPAS_PERSIST_WPE > 0
#define MPAS_PERSIST_WAVES __attribute__((amdgpu_waves_per_eu(MPAS_PERSIST_WPE, 8)))
#else
#define MPAS_PERSIST_WAVES
#endif
struct PersistArgs {
  const BatchGroup *__restrict__ grp;
  const int4 *__restrict__ tab[kBatchKernels];  // (mode, group, block, 0)
  int n[kBatchKernels];                          // entries of each table
  int xcd;                                       // XCD-contiguous walk
};

template <int NT, int W, int WE, int SCH>
__device__ __forceinline__ void run_phase(int k, int mode, const BatchGroup &g, int s, int item) {
#define MPAS_MODES(call)                  \
  switch (mode) {                         \
    case 1: { constexpr int MODE = 1; call; } break; \
    case 2: { constexpr int MODE = 2; call; } break; \
    case 3: { constexpr int MODE = 3; call; } break; \
    case 4: { constexpr int MODE = 4; call; } break; \
    case 5: { constexpr int MODE = 5; call; } break; \
    default: { constexpr int MODE = 6; call; } break; \
  }
  switch (k) {
    case 0: MPAS_MODES((run_edges<WE, MODE>(g, s, item))) break;
    case 1:
      if constexpr (NT > 0) MPAS_MODES((run_ho<NT, WE, MODE>(g, s, item)))
      break;
    case 2: MPAS_MODES((run_cells<NT, W, MODE, SCH>(g, s, item))) break;
    case 3:
      if constexpr (NT > 0) MPAS_MODES((run_tracer<NT, W, MODE>(g, s, item)))
      break;
    case 4: MPAS_MODES((diag_vertex_lanes<MODE>(g, s, item))) break;
    default: MPAS_MODES((diag_edge_lanes<MODE>(g, s, item))) break;
  }
#undef MPAS_MODES
}

// Waves per SIMD the persistent kernel is compiled for (its registers are
// the largest body's; 0: the compiler's choice).
#ifndef MPAS_PERSIST_WPE
#define MPAS_PERSIST_WPE 2
#endif
template <int NT, int W, int WE, int SCH>
__global__ __launch_bounds__(kPBlock) MPAS_PERSIST_WAVES void kb_persistent(PersistArgs P,
                                                                           int nsteps) {
  namespace cg = cooperative_groups;
  cg::grid_group grid = cg::this_grid();
  // XCD x (workgroups b with b % 8 == x) walks one contiguous range of each
  // table, so neighbouring blocks share that XCD's L2
  const int nb = int(gridDim.x), b = int(blockIdx.x);
  const int per_xcd = nb / kXcds, x = b % kXcds, j = b / kXcds;
  for (int step = 0; step < nsteps; step++)
    for (int s = 0; s < 4; s++)
      for (int k = 0; k < kBatchKernels; k++) {
        const int n = P.n[k];
        if (!n) continue;
        int hi = n, first = b, stride = nb;
        if (P.xcd && per_xcd > 0) {
          if (b < per_xcd * kXcds) {
            const int chunk = (n + kXcds - 1) / kXcds;
            first = x * chunk + j;
            hi = min(n, x * chunk + chunk);
            stride = per_xcd;
          } else {
            first = n;  // workgroups past whole rounds of 8 stay idle
          }
        }
        for (int lb = first; lb < hi; lb += stride) {
          const int4 t = P.tab[k][lb];
          if (t.y < 0) continue;
          run_phase<NT, W, WE, SCH>(k, t.x, P.grp[t.y], s, t.z * kPBlock + int(threadIdx.x));
        }
        grid.sync();
      }
}

// One member's field between its contiguous copy (n entities x w doubles)
// and the group layout (n entities x members x w doubles).
// Entity e of the group's (possibly reordered) mesh is entity of[e] of the
// caller's numbering (of = nullptr: the same).
__global__ __launch_bounds__(256) void kb_member_scatter(double *group, const double *src,
                                                         int64_t n, int w, int members, int idx,
                                                         const int32_t *of) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * w) return;
  const int64_t e = i / w, j = i - e * w;
  const int64_t o = of ? int64_t(of[e]) : e;
  group[(e * members + idx) * w + j] = src[o * w + j];
}
__global__ __launch_bounds__(256) void kb_member_gather(double *dst, const double *group,
                                                        int64_t n, int w, int members, int idx,
                                                        const int32_t *of) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * w) return;
  const int64_t e = i / w, j = i - e * w;
  const int64_t o = of ? int64_t(of[e]) : e;
  dst[o * w + j] = group[(e * members + idx) * w + j];
}
// Rows of r doubles from one entity order to another: dst row i = src row
// map[i].
__global__ __launch_bounds__(256) void kb_permute_rows(double *dst, const double *src, int64_t n,
                                                       int64_t r, const int32_t *map) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n * r) return;
  const int64_t e = i / r, j = i - e * r;
  dst[i] = src[int64_t(map[e]) * r + j];
}

// The context group q's mesh runs on in the batch's mesh order.
mpas_ctx *mesh_ctx(const mpas_batch *B, int mesh) {
  return B->mesh_order ? B->octx[B->mesh_order][mesh] : B->meshes[mesh];
}

int lanes_per_entity(int nT) { return nT > 0 ? nT / (nT % 2 == 0 ? 2 : 1) : 1; }

// Whether group q runs one entity per wave when each lane takes `per_lane`
// times the members it takes otherwise: when the entity's lanes fill at
// least three quarters of their waves.
bool group_wave_entity(const mpas_batch *B, int q, int per_lane) {
  if (B->wave_entity != 1) return B->wave_entity == 2;
  const int64_t nch = (int64_t(B->groups[q].members.size()) + B->lpl - 1) / B->lpl / per_lane;
  const int64_t n = nch * lanes_per_entity(B->nT);
  return 4 * n >= 3 * ((n + 63) / 64 * 64);
}

// The layout of group q in kernel k (0 momentum, 1 high-order flux, 2
// cells, 3 tracer update, 4 vertex and 5 edge diagnostics): levels on
// lanes (odd) or one entity per wave (even); with one member per lane
// 1 and 2, and for the kernels in level-group form 3 and 4 (two members
// per lane) when the group's member count is even; with more members per
// lane 5 and 6.
int group_mode(const mpas_batch *B, int q, int k) {
  if (B->lpl > 1) return group_wave_entity(B, q, 1) ? 6 : 5;
  const bool grouped = k == 0 || k == 1 || k == 4 || k == 5;
  if (grouped && MPAS_LEVEL_PAIRS && B->groups[q].members.size() % 2 == 0)
    return group_wave_entity(B, q, 2) ? 4 : 3;
  return group_wave_entity(B, q, 1) ? 2 : 1;
}

// Lanes per entity of group q in kernel k (tracer lane groups npl when it
// has them): nchunk * npl (half the chunks with level pairs), padded to a
// multiple of the wave size with one entity per wave.
int64_t entity_lanes(const mpas_batch *B, int q, int k) {
  const GroupInfo &G = B->groups[q];
  const bool tracer_lanes = k >= 1 && k <= 3;
  const int64_t npl = tracer_lanes ? lanes_per_entity(B->nT) : 1;
  int64_t nch = (int64_t(G.members.size()) + B->lpl - 1) / B->lpl;
  const int mode = group_mode(B, q, k);
  if (lp_of(mode) == 2) nch /= 2;
  const int64_t n = nch * npl;
  return mode % 2 == 0 ? (n + 63) / 64 * 64 : n;
}

// Work items of kernel k for group q.
int64_t kernel_items(const mpas_batch *B, int k, int q) {
  const DMesh &M = mesh_ctx(B, B->groups[q].ctx)->dm;
  switch (k) {
    case 0: return M.nE * entity_lanes(B, q, k);
    case 1: return B->nT ? M.nE * entity_lanes(B, q, k) : 0;
    case 2: return M.nC * entity_lanes(B, q, k);
    case 3: return B->nT ? M.nC * entity_lanes(B, q, k) : 0;
    case 4: return M.nV * entity_lanes(B, q, k);
    default: return M.nE * entity_lanes(B, q, k);
  }
}

// Workgroup size of kernel k in layout mode.
int block_of(const mpas_batch *B, int k, int mode) {
  const int b = B->mblock[k][mode - 1];
  return b ? b : B->block[k];
}

void free_tables(mpas_batch *B) {
  for (int k = 0; k < kBatchKernels; k++) {
    for (int m = 0; m < kBatchModes; m++) {
      if (B->d_tab[k][m]) (void)hipFree(B->d_tab[k][m]);
      B->d_tab[k][m] = nullptr;
      B->grid[k][m] = 0;
    }
    if (B->d_ptab[k]) (void)hipFree(B->d_ptab[k]);
    B->d_ptab[k] = nullptr;
    B->pn[k] = 0;
  }
  B->tables_ok = false;
}

void drop_graph(mpas_batch *B) {
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
}

// The stage descriptors of a group: the roles sw_rk4_steps gives the
// context's buffers, with the ping-pong slot of stage s = s % 2.
void describe_group(mpas_batch *B, int q) {
  const GroupInfo &G = B->groups[q];
  const GroupBufs &W = G.b;
  mpas_ctx *ctx = mesh_ctx(B, G.ctx);
  BatchGroup &d = B->desc[q];
  d.M = ctx->dm;
  d.M.nL = int(G.members.size());
  d.pv_vertex = W.pv_vertex;
  for (int s = 0; s < 4; s++) {
    const int cur = s % 2, nxt = 1 - cur;
    StageIO S = stage_params(ctx, s, G.dt);
    S.lpl = B->lpl;
    S.nchunk = (d.M.nL + B->lpl - 1) / B->lpl;
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
    S.keep_rfac = 0;
    S.tlo = W.tlo;
    d.S[s] = S;
    d.vort_y[s] = W.vort[nxt];
    d.h_edge_y[s] = W.h_edge[nxt];
    d.pv_edge_y[s] = W.pv_edge[nxt];
  }
}

// Descriptors (they hold members per lane) and workgroup tables.
int build_tables(mpas_batch *B) {
  drop_graph(B);
  free_tables(B);
  for (size_t q = 0; q < B->groups.size(); q++) describe_group(B, int(q));
  MPAS_HIP(hipMemcpy(B->d_grp, B->desc.data(), sizeof(BatchGroup) * B->desc.size(),
                     hipMemcpyHostToDevice));
  const int ng = int(B->groups.size());
  for (int k = 0; k < kBatchKernels; k++)
    for (int mode = 1; mode <= kBatchModes; mode++) {
      const int bs = block_of(B, k, mode);
      std::vector<int> nb(ng, 0);
      for (int q = 0; q < ng; q++)
        if (group_mode(B, q, k) == mode) nb[q] = int((kernel_items(B, k, q) + bs - 1) / bs);
      std::vector<int2> tab;
      if (B->order == 1) {  // round-robin over the groups' blocks
        for (int j = 0, more = 1; more; j++) {
          more = 0;
          for (int q = 0; q < ng; q++)
            if (j < nb[q]) tab.push_back(make_int2(q, j)), more = 1;
        }
      } else {
        for (int q = 0; q < ng; q++)
          for (int j = 0; j < nb[q]; j++) tab.push_back(make_int2(q, j));
      }
      if (tab.empty()) continue;
      while (tab.size() % kXcds) tab.push_back(make_int2(-1, 0));
      B->grid[k][mode - 1] = int(tab.size());
      MPAS_HIP(hipMalloc(&B->d_tab[k][mode - 1], sizeof(int2) * tab.size()));
      MPAS_HIP(hipMemcpy(B->d_tab[k][mode - 1], tab.data(), sizeof(int2) * tab.size(),
                         hipMemcpyHostToDevice));
    }
  // the persistent launch: every layout of kernel k in one table, blocks
  // of kPBlock items, each group's blocks together
  for (int k = 0; k < kBatchKernels; k++) {
    std::vector<int4> tab;
    for (int mode = 1; mode <= kBatchModes; mode++)
      for (int q = 0; q < ng; q++)
        if (group_mode(B, q, k) == mode) {
          const int nb = int((kernel_items(B, k, q) + kPBlock - 1) / kPBlock);
          for (int j = 0; j < nb; j++) tab.push_back(make_int4(mode, q, j, 0));
        }
    if (tab.empty()) continue;
    B->pn[k] = int(tab.size());
    MPAS_HIP(hipMalloc(&B->d_ptab[k], sizeof(int4) * tab.size()));
    MPAS_HIP(hipMemcpy(B->d_ptab[k], tab.data(), sizeof(int4) * tab.size(),
                       hipMemcpyHostToDevice));
  }
  B->tables_ok = true;
  return 0;
}

BatchArgs args_of(const mpas_batch *B, int k, int mode) {
  return BatchArgs{B->d_grp, B->d_tab[k][mode - 1], B->xcd};
}

// One phase: the launch of the groups of each layout.
#define BL1(k, kern, m)                                                                   \
  if (B->grid[k][m - 1]) {                                                                \
    constexpr int MODE = m;                                                               \
    hipLaunchKernelGGL(kern, dim3(B->grid[k][m - 1]), dim3(block_of(B, k, m)), 0, st,      \
                       args_of(B, k, m), s);                                              \
    if (int rc = after_launch(#kern)) return rc;                                          \
  }
#define BL(k, kern)                                                                       \
  do {                                                                                    \
    BL1(k, kern, 1) BL1(k, kern, 2) BL1(k, kern, 3) BL1(k, kern, 4) BL1(k, kern, 5)       \
    BL1(k, kern, 6)                                                                       \
  } while (0)
static_assert(kBatchModes == 6, "BL launches modes 1-6");

template <int NT, int W, int WE>
int launch_batch_stage_t(mpas_batch *B, int s) {
  hipStream_t st = B->stream;
  const bool fct = NT > 0 && B->desc[0].S[0].scheme == MPAS_TRACER_FCT;
  BL(0, (kb_stage_edges<WE, MODE>));
  if constexpr (NT > 0) {
    if (fct) BL(1, (kb_stage_ho<NT, WE, MODE>));
  }
  // the tracer scheme is fixed per instantiation (one set of registers each)
  const int sch = B->desc[0].S[0].scheme;
  if (NT == 0 || sch == MPAS_TRACER_NONE)
    BL(2, (kb_stage_cells<NT, W, MODE, MPAS_TRACER_NONE>));
  else if (sch == MPAS_TRACER_CENTRED)
    BL(2, (kb_stage_cells<NT, W, MODE, MPAS_TRACER_CENTRED>));
  else
    BL(2, (kb_stage_cells<NT, W, MODE, MPAS_TRACER_FCT>));
  if constexpr (NT > 0) {
    if (fct) BL(3, (kb_stage_tracer<NT, W, MODE>));
  }
  BL(4, kb_diag_vertex<MODE>);
  BL(5, kb_diag_edge<MODE>);
  return 0;
}
#undef BL
#undef BL1

template <int NT>
int launch_batch_stage_w(mpas_batch *B, int s) {
  if (B->narrow) return launch_batch_stage_t<NT, 6, 10>(B, s);
  return launch_batch_stage_t<NT, 8, 16>(B, s);
}

int launch_batch_stage(mpas_batch *B, int s) {
  switch (B->nT) {
    case 0: return launch_batch_stage_w<0>(B, s);
    case 1: return launch_batch_stage_w<1>(B, s);
    case 2: return launch_batch_stage_w<2>(B, s);
    case 3: return launch_batch_stage_w<3>(B, s);
    case 4: return launch_batch_stage_w<4>(B, s);
    case 5: return launch_batch_stage_w<5>(B, s);
    case 6: return launch_batch_stage_w<6>(B, s);
    case 7: return launch_batch_stage_w<7>(B, s);
    case 8: return launch_batch_stage_w<8>(B, s);
  }
  MPAS_FAIL(MPAS_EINVAL, "ensembles support at most 8 tracers");
}

int launch_batch_step(mpas_batch *B) {
  for (int s = 0; s < 4; s++)
    if (int rc = launch_batch_stage(B, s)) return rc;
  return 0;
}

// The persistent launch of nsteps steps (see kb_persistent).
template <int NT, int SCH>
int launch_persistent_t(mpas_batch *B, int nsteps) {
  const auto kern = kb_persistent<NT, 6, 10, SCH>;
  int per_cu = 0;
  MPAS_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kPBlock, 0));
  if (per_cu < 1) MPAS_FAIL(MPAS_EHIP, "the persistent ensemble kernel does not fit a CU");
  int grid = per_cu * B->cus;
  if (B->xcd) grid = std::max(kXcds, grid / kXcds * kXcds);
  PersistArgs P{};
  P.grp = B->d_grp;
  for (int k = 0; k < kBatchKernels; k++) {
    P.tab[k] = B->d_ptab[k];
    P.n[k] = B->pn[k];
  }
  P.xcd = B->xcd;
  int n = nsteps;
  void *args[] = {&P, &n};
  MPAS_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(kern), dim3(grid),
                                      dim3(kPBlock), args, 0, B->stream));
  return 0;
}

// Whether the persistent launch is built for this ensemble's shapes.
bool persistent_supported(const mpas_batch *B) {
  const int sch = B->desc.empty() ? MPAS_TRACER_NONE : B->desc[0].S[0].scheme;
  return B->narrow && B->cus > 0 &&
         ((B->nT == 0) || ((B->nT == 1 || B->nT == 2 || B->nT == 4) && sch == MPAS_TRACER_FCT));
}

int launch_persistent(mpas_batch *B, int nsteps) {
  if (!persistent_supported(B))
    MPAS_FAIL(MPAS_EINVAL, "the persistent launch serves meshes of at most 6 edges per cell "
                           "with 0, 1, 2 or 4 FCT tracers (use MPAS_BATCH_GRAPH)");
  switch (B->nT) {
    case 0: return launch_persistent_t<0, MPAS_TRACER_NONE>(B, nsteps);
    case 1: return launch_persistent_t<1, MPAS_TRACER_FCT>(B, nsteps);
    case 2: return launch_persistent_t<2, MPAS_TRACER_FCT>(B, nsteps);
    default: return launch_persistent_t<4, MPAS_TRACER_FCT>(B, nsteps);
  }
}

int capture_graph(mpas_batch *B) {
  drop_graph(B);
  MPAS_HIP(hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal));
  const int rc = launch_batch_step(B);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(B->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess)
    MPAS_FAIL(MPAS_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  B->graph = g;
  MPAS_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
  return 0;
}

// Carves one group's buffers out of the arena (256-byte aligned).
struct Carver {
  char *base;
  size_t off = 0;
  double *take(size_t n) {
    double *p = base ? reinterpret_cast<double *>(base + off) : nullptr;
    off += (n * sizeof(double) + 255) / 256 * 256;
    return p;
  }
};

void carve_group(Carver &c, const DMesh &M, int members, int nT, GroupBufs &b) {
  const size_t E = size_t(M.nE) * members, C = size_t(M.nC) * members;
  const size_t V = size_t(M.nV) * members, T = C * nT;
  b.u = c.take(E), b.nu = c.take(E), b.pu[0] = c.take(E), b.pu[1] = c.take(E);
  b.h_edge[0] = c.take(E), b.h_edge[1] = c.take(E), b.pv_edge[0] = c.take(E), b.pv_edge[1] = c.take(E);
  b.h = c.take(C), b.nh = c.take(C), b.ph[0] = c.take(C), b.ph[1] = c.take(C);
  b.ke[0] = c.take(C), b.ke[1] = c.take(C), b.div[0] = c.take(C), b.div[1] = c.take(C);
  b.h_s = c.take(size_t(M.nC));
  b.tr = c.take(T), b.ntr = c.take(T), b.ptr[0] = c.take(T), b.ptr[1] = c.take(T), b.tlo = c.take(T);
  b.vort[0] = c.take(V), b.vort[1] = c.take(V), b.pv_vertex = c.take(V);
  b.anti = c.take(E * nT), b.rfac = c.take(2 * T);
}

int check_batch(const mpas_batch *B) {
  if (!B) MPAS_FAIL(MPAS_EINVAL, "batch is NULL");
  MPAS_HIP(hipSetDevice(B->device));
  return 0;
}

void release_batch(mpas_batch *B) {
  drop_graph(B);
  free_tables(B);
  for (int o = 1; o < 3; o++) {
    for (mpas_ctx *c : B->octx[o])
      if (c) (void)mpas_ctx_destroy(c);
    for (int32_t *p : B->d_cell_of[o])
      if (p) (void)hipFree(p);
    for (int32_t *p : B->d_edge_of[o])
      if (p) (void)hipFree(p);
  }
  if (B->d_grp) (void)hipFree(B->d_grp);
  if (B->staging) (void)hipFree(B->staging);
  if (B->arena) (void)hipFree(B->arena);
  if (B->stream) (void)hipStreamDestroy(B->stream);
}

int run_steps(mpas_batch *B, int nsteps, int mode) {
  if (!B->tables_ok)
    if (int rc = build_tables(B)) return rc;
  if (mode == MPAS_BATCH_TUNED) mode = B->tuned_mode;
  if (mode == MPAS_BATCH_PERSISTENT) {
    // launches of at most 20 steps (each some tens of milliseconds)
    for (int done = 0; done < nsteps;) {
      const int n = std::min(20, nsteps - done);
      if (int rc = launch_persistent(B, n)) return rc;
      done += n;
    }
    return 0;
  }
  if (mode == MPAS_BATCH_LAUNCH) {
    for (int step = 0; step < nsteps; step++)
      if (int rc = launch_batch_step(B)) return rc;
    return 0;
  }
  if (!B->exec)
    if (int rc = capture_graph(B)) return rc;
  for (int step = 0; step < nsteps; step++) MPAS_HIP(hipGraphLaunch(B->exec, B->stream));
  return 0;
}

double time_steps(mpas_batch *B, int nsteps, int *rc, int mode = MPAS_BATCH_GRAPH) {
  *rc = run_steps(B, 1, mode);  // capture + warm
  if (!*rc && hipStreamSynchronize(B->stream) != hipSuccess) *rc = MPAS_EHIP;
  if (*rc) return 0;
  const auto t0 = std::chrono::steady_clock::now();
  *rc = run_steps(B, nsteps, mode);
  if (!*rc && hipStreamSynchronize(B->stream) != hipSuccess) *rc = MPAS_EHIP;
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// One member's field between its host array (n entities x w doubles) and
// the group layout, through the staging buffer: one contiguous copy and a
// scatter / gather kernel.
// `cells`: the field is per cell (else per edge), for the mesh order.
int member_copy(mpas_batch *B, int m, double *group_field, const double *host, double *host_out,
                size_t n, size_t w, bool cells = true) {
  const GroupInfo &G = B->groups[B->grp_of[m]];
  const int members = int(G.members.size()), i = B->idx_in_grp[m];
  const int o = B->mesh_order;
  const int32_t *of = o ? (cells ? B->d_cell_of[o][G.ctx] : B->d_edge_of[o][G.ctx]) : nullptr;
  const int64_t cnt = int64_t(n * w);
  const int grid = int((cnt + 255) / 256);
  if (host) {
    MPAS_HIP(hipMemcpyAsync(B->staging, host, sizeof(double) * cnt, hipMemcpyHostToDevice,
                            B->stream));
    hipLaunchKernelGGL(kb_member_scatter, dim3(grid), dim3(256), 0, B->stream, group_field,
                       B->staging, int64_t(n), int(w), members, i, of);
    return after_launch("kb_member_scatter");
  }
  hipLaunchKernelGGL(kb_member_gather, dim3(grid), dim3(256), 0, B->stream, B->staging,
                     group_field, int64_t(n), int(w), members, i, of);
  if (int rc = after_launch("kb_member_gather")) return rc;
  MPAS_HIP(hipMemcpyAsync(host_out, B->staging, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                          B->stream));
  MPAS_HIP(hipStreamSynchronize(B->stream));
  return 0;
}

// The reordered contexts of order o and their maps to the original ids.
int ensure_order(mpas_batch *B, int o) {
  if (o == 0 || !B->octx[o].empty()) return 0;
  const size_t nm = B->meshes.size();
  B->octx[o].assign(nm, nullptr);
  B->d_cell_of[o].assign(nm, nullptr);
  B->d_edge_of[o].assign(nm, nullptr);
  for (size_t q = 0; q < nm; q++) {
    if (int rc = mpas_ctx_create_reordered(B->meshes[q], o, &B->octx[o][q])) return rc;
    const mpas_ctx *c = B->octx[o][q];
    const HostMesh &h = c->host;
    MPAS_HIP(hipMalloc(&B->d_cell_of[o][q], sizeof(int32_t) * h.nC));
    MPAS_HIP(hipMalloc(&B->d_edge_of[o][q], sizeof(int32_t) * h.nE));
    MPAS_HIP(hipMemcpy(B->d_cell_of[o][q], h.cellOf.data(), sizeof(int32_t) * h.nC,
                       hipMemcpyHostToDevice));
    MPAS_HIP(hipMemcpy(B->d_edge_of[o][q], h.edgeOf.data(), sizeof(int32_t) * h.nE,
                       hipMemcpyHostToDevice));
  }
  return 0;
}

// Original ids of a mesh's cells or edges in order o (identity for 0).
std::vector<int32_t> original_ids(const mpas_batch *B, int o, int mesh, bool cells) {
  const mpas_ctx *c = o ? B->octx[o][mesh] : B->meshes[mesh];
  const HostMesh &h = c->host;
  const int n = cells ? h.nC : h.nE;
  const std::vector<int32_t> &of = cells ? h.cellOf : h.edgeOf;
  std::vector<int32_t> r(n);
  for (int i = 0; i < n; i++) r[i] = of.empty() ? i : of[i];
  return r;
}

// The groups' diagnostics of their current state.
int group_diagnostics(mpas_batch *B) {
  for (size_t q = 0; q < B->groups.size(); q++) {
    GroupInfo &G = B->groups[q];
    const mpas_ctx *ctx = mesh_ctx(B, G.ctx);
    DMesh Mg = ctx->dm;
    Mg.nL = int(G.members.size());
    if (int rc = diagnostics_dm(Mg, B->stream, ctx->params.visc2 != 0.0, G.b.u, G.b.h,
                                G.b.ke[0], G.b.div[0], G.b.vort[0], G.b.pv_vertex,
                                G.b.h_edge[0], G.b.pv_edge[0]))
      return rc;
  }
  return 0;
}

// The members' state (u, h, tracers, h_s), held in order `from` in the arena
// image `src`, into the arena in order `to`, then the diagnostics there.
int permute_state(mpas_batch *B, const char *src, int from, int to) {
  if (int rc = ensure_order(B, to)) return rc;
  const char *arena = static_cast<const char *>(B->arena);
  std::vector<int32_t *> maps;
  int rc = 0;
  for (size_t q = 0; q < B->groups.size() && !rc; q++) {
    GroupInfo &G = B->groups[q];
    const int members = int(G.members.size());
    for (int cells = 0; cells < 2 && !rc; cells++) {
      const std::vector<int32_t> of_from = original_ids(B, from, G.ctx, cells);
      const std::vector<int32_t> of_to = original_ids(B, to, G.ctx, cells);
      const int n = int(of_to.size());
      std::vector<int32_t> pos(n), map(n);
      for (int i = 0; i < n; i++) pos[of_from[i]] = i;
      for (int i = 0; i < n; i++) map[i] = pos[of_to[i]];
      int32_t *d = nullptr;
      if (hipMalloc(&d, sizeof(int32_t) * n) != hipSuccess) {
        set_error("mpas_batch: no memory for a mesh order map");
        rc = MPAS_ENOMEM;
        break;
      }
      maps.push_back(d);
      if (hipMemcpy(d, map.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice) != hipSuccess) {
        set_error("mpas_batch: copy of a mesh order map failed");
        rc = MPAS_EHIP;
        break;
      }
      struct F { double *f; int64_t r; };
      std::vector<F> fields;
      if (cells) {
        fields = {{G.b.h, members}, {G.b.h_s, 1}};
        if (B->nT) fields.push_back({G.b.tr, int64_t(members) * B->nT});
      } else {
        fields = {{G.b.u, members}};
      }
      for (const F &f : fields) {
        const double *s_ = reinterpret_cast<const double *>(
            src + (reinterpret_cast<const char *>(f.f) - arena));
        const int64_t cnt = int64_t(n) * f.r;
        hipLaunchKernelGGL(kb_permute_rows, dim3(int((cnt + 255) / 256)), dim3(256), 0,
                           B->stream, f.f, s_, int64_t(n), f.r, d);
        if ((rc = after_launch("kb_permute_rows"))) break;
      }
    }
  }
  if (!rc) {
    B->mesh_order = to;
    if (!(rc = build_tables(B))) rc = group_diagnostics(B);
  }
  (void)hipStreamSynchronize(B->stream);
  for (int32_t *d : maps) (void)hipFree(d);
  return rc;
}

}  // namespace
}  // namespace mpasx

using namespace mpasx;

extern "C" {

int mpas_batch_create(mpas_ctx *const *meshes, int nmeshes, int nmembers, const int32_t *mesh_of,
                      const double *dt, mpas_batch **out) {
  if (!out) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_create: out is NULL");
  *out = nullptr;
  if (!meshes || nmeshes < 1 || nmembers < 1 || !mesh_of || !dt)
    MPAS_FAIL(MPAS_EINVAL, "mpas_batch_create: bad arguments");
  auto B = std::make_unique<mpas_batch>();
  for (int q = 0; q < nmeshes; q++) {
    if (int rc = check_ctx(meshes[q])) return rc;
    B->meshes.push_back(meshes[q]);
  }
  const mpas_ctx *c0 = meshes[0];
  B->device = c0->device;
  {
    int coop = 0;
    MPAS_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, B->device));
    if (coop)
      MPAS_HIP(hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, B->device));
  }
  B->nT = c0->dm.nT;
  B->narrow = c0->dm.narrow;
  for (mpas_ctx *c : B->meshes) {
    if (c->device != B->device) MPAS_FAIL(MPAS_EINVAL, "every mesh context must use one device");
    if (c->dm.nT != B->nT || c->params.tracer_scheme != c0->params.tracer_scheme)
      MPAS_FAIL(MPAS_EINVAL, "every mesh context must have the same tracers and scheme");
    if (c->dm.narrow != B->narrow)
      MPAS_FAIL(MPAS_EINVAL, "mesh contexts mix narrow and wide stencil widths");
    if (c->dm.nL != 1)
      MPAS_FAIL(MPAS_EINVAL, "ensemble mesh contexts have one level (members take the levels)");
    if (B->nT && c->params.tracer_scheme == MPAS_TRACER_FCT &&
        (c->dm.mlc || c->dm.hoMask || !c->aligned_c2c))
      MPAS_FAIL(MPAS_EINVAL, "the fused FCT step needs every level active and aligned cellsOnCell");
  }
  B->nmem = nmembers;
  B->grp_of.resize(nmembers);
  B->idx_in_grp.resize(nmembers);
  for (int m = 0; m < nmembers; m++) {
    if (mesh_of[m] < 0 || mesh_of[m] >= nmeshes)
      MPAS_FAIL(MPAS_EINVAL, "mpas_batch_create: mesh_of out of range");
    if (!(dt[m] > 0)) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_create: dt must be positive");
    int q = 0;
    for (; q < int(B->groups.size()); q++)
      if (B->groups[q].ctx == mesh_of[m] && B->groups[q].dt == dt[m]) break;
    if (q == int(B->groups.size())) {
      B->groups.emplace_back();
      B->groups[q].ctx = mesh_of[m];
      B->groups[q].dt = dt[m];
    }
    B->grp_of[m] = q;
    B->idx_in_grp[m] = int(B->groups[q].members.size());
    B->groups[q].members.push_back(m);
  }
  // the kernels address each group array with 32-bit byte offsets; the
  // largest are the antidiffusive fluxes (edges x members x tracers), the
  // limiter factors (2 x cells x members x tracers) and the edge fields
  for (const GroupInfo &G : B->groups) {
    const DMesh &M = B->meshes[G.ctx]->dm;
    const size_t m = G.members.size(), nt = size_t(std::max(B->nT, 1));
    const size_t largest = std::max({size_t(M.nE) * m * nt, 2 * size_t(M.nC) * m * nt,
                                     size_t(M.nE) * m});
    if (largest > size_t(UINT32_MAX) / sizeof(double))
      MPAS_FAIL(MPAS_ESHAPE, "mpas_batch_create: " + std::to_string(m) +
                                 " members of one mesh and time step make a " +
                                 std::to_string(largest * sizeof(double)) +
                                 "-byte array; the kernels address at most 4 GiB per array");
  }
  Carver sizer{nullptr};
  for (GroupInfo &G : B->groups) {
    GroupBufs dummy;
    carve_group(sizer, B->meshes[G.ctx]->dm, int(G.members.size()), B->nT, dummy);
  }
  B->arena_bytes = sizer.off;
  if (hipMalloc(&B->arena, B->arena_bytes) != hipSuccess) {
    release_batch(B.get());
    MPAS_FAIL(MPAS_ENOMEM, "hipMalloc of " + std::to_string(B->arena_bytes) +
                               " bytes for the ensemble failed");
  }
  if (zero_now(B->arena, B->arena_bytes) != hipSuccess ||
      hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&B->d_grp, sizeof(BatchGroup) * B->groups.size()) != hipSuccess) {
    release_batch(B.get());
    MPAS_FAIL(MPAS_EHIP, "ensemble set-up failed");
  }
  size_t stage_n = 1;
  for (const GroupInfo &G : B->groups) {
    const DMesh &M = B->meshes[G.ctx]->dm;
    stage_n = std::max({stage_n, size_t(M.nE), size_t(M.nC) * std::max(1, B->nT)});
  }
  if (hipMalloc(&B->staging, sizeof(double) * stage_n) != hipSuccess) {
    release_batch(B.get());
    MPAS_FAIL(MPAS_EHIP, "ensemble set-up failed");
  }
  Carver c{static_cast<char *>(B->arena)};
  for (GroupInfo &G : B->groups)
    carve_group(c, B->meshes[G.ctx]->dm, int(G.members.size()), B->nT, G.b);
  B->desc.resize(B->groups.size());
  B->ready.assign(nmembers, 0);
  if (int rc = build_tables(B.get())) {
    release_batch(B.get());
    return rc;
  }
  *out = B.release();
  return 0;
}

int mpas_batch_destroy(mpas_batch *B) {
  if (!B) return 0;
  (void)hipSetDevice(B->device);
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  release_batch(B);
  delete B;
  return 0;
}

int mpas_batch_upload(mpas_batch *B, int member, const double *h, const double *u,
                      const double *tracers, const double *h_s) {
  if (int rc = check_batch(B)) return rc;
  if (member < 0 || member >= B->nmem) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_upload: bad member");
  if (!h || !u || !h_s || (B->nT && !tracers))
    MPAS_FAIL(MPAS_EINVAL, "mpas_batch_upload: NULL array");
  GroupInfo &G = B->groups[B->grp_of[member]];
  const mpas_ctx *ctx = mesh_ctx(B, G.ctx);
  const DMesh &M = ctx->dm;
  {
    int rc;
    if ((rc = check_finite(h, size_t(M.nC), "mpas_batch_upload: h")) ||
        (rc = check_finite(u, size_t(M.nE), "mpas_batch_upload: u")) ||
        (rc = check_finite(h_s, size_t(M.nC), "mpas_batch_upload: h_s")) ||
        (B->nT && (rc = check_finite(tracers, size_t(M.nC) * B->nT, "mpas_batch_upload: tracers"))))
      return rc;
  }
  if (G.h_s.empty()) {
    G.h_s.assign(h_s, h_s + M.nC);
    const int32_t *of = B->mesh_order ? B->d_cell_of[B->mesh_order][G.ctx] : nullptr;
    MPAS_HIP(hipMemcpyAsync(B->staging, h_s, sizeof(double) * M.nC, hipMemcpyHostToDevice,
                            B->stream));
    hipLaunchKernelGGL(kb_member_scatter, dim3((M.nC + 255) / 256), dim3(256), 0, B->stream,
                       G.b.h_s, B->staging, int64_t(M.nC), 1, 1, 0, of);
    if (int rc = after_launch("kb_member_scatter")) return rc;
  } else if (std::memcmp(G.h_s.data(), h_s, sizeof(double) * M.nC) != 0) {
    MPAS_FAIL(MPAS_EINVAL, "members of one mesh and time step share h_s (bottom topography)");
  }
  int rc;
  if ((rc = member_copy(B, member, G.b.h, h, nullptr, M.nC, 1)) ||
      (rc = member_copy(B, member, G.b.u, u, nullptr, M.nE, 1, false)) ||
      (B->nT && (rc = member_copy(B, member, G.b.tr, tracers, nullptr, M.nC, B->nT))))
    return rc;
  // diagnostics of the group's state (every member: cheap, and the members
  // not yet uploaded are recomputed when they are)
  DMesh Mg = M;
  Mg.nL = int(G.members.size());
  if ((rc = diagnostics_dm(Mg, B->stream, ctx->params.visc2 != 0.0, G.b.u, G.b.h, G.b.ke[0],
                           G.b.div[0], G.b.vort[0], G.b.pv_vertex, G.b.h_edge[0],
                           G.b.pv_edge[0])))
    return rc;
  MPAS_HIP(hipStreamSynchronize(B->stream));
  B->ready[member] = 1;
  return 0;
}

int mpas_batch_download(mpas_batch *B, int member, double *h, double *u, double *tracers) {
  if (int rc = check_batch(B)) return rc;
  if (member < 0 || member >= B->nmem) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_download: bad member");
  GroupInfo &G = B->groups[B->grp_of[member]];
  const DMesh &M = B->meshes[G.ctx]->dm;
  int rc;
  if ((h && (rc = member_copy(B, member, G.b.h, nullptr, h, M.nC, 1))) ||
      (u && (rc = member_copy(B, member, G.b.u, nullptr, u, M.nE, 1, false))) ||
      (tracers && B->nT && (rc = member_copy(B, member, G.b.tr, nullptr, tracers, M.nC, B->nT))))
    return rc;
  MPAS_HIP(hipStreamSynchronize(B->stream));
  return 0;
}

int mpas_batch_step(mpas_batch *B, int nsteps, int mode) {
  if (int rc = check_batch(B)) return rc;
  if (mode < MPAS_BATCH_GRAPH || mode > MPAS_BATCH_TUNED)
    MPAS_FAIL(MPAS_EINVAL, "mpas_batch_step: unknown mode");
  for (int m = 0; m < B->nmem; m++)
    if (!B->ready[m]) MPAS_FAIL(MPAS_ESTATE, "member " + std::to_string(m) + " has no state");
  if (nsteps < 0) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_step: nsteps < 0");
  return run_steps(B, nsteps, mode);
}

int mpas_batch_sync(mpas_batch *B) {
  if (int rc = check_batch(B)) return rc;
  MPAS_HIP(hipStreamSynchronize(B->stream));
  return 0;
}

int mpas_batch_groups(const mpas_batch *B) { return B ? int(B->groups.size()) : MPAS_EINVAL; }

int mpas_batch_config(mpas_batch *B, const int32_t *block, int members_per_lane, int order,
                      int xcd, int wave_entity) {
  if (int rc = check_batch(B)) return rc;
  if (order < 0 || order > 1) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_config: order is 0 or 1");
  if (members_per_lane < 1) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_config: members_per_lane < 1");
  if (block)
    for (int k = 0; k < kBatchKernels; k++)
      if (block[k] < 64 || block[k] > kBatchMaxBlock || block[k] % 64)
        MPAS_FAIL(MPAS_EINVAL, "mpas_batch_config: block sizes are multiples of 64 up to 512");
  MPAS_HIP(hipStreamSynchronize(B->stream));
  if (block) std::copy(block, block + kBatchKernels, B->block);
  for (auto &row : B->mblock)  // an explicit configuration applies to every layout
    for (int &b : row) b = 0;
  B->lpl = members_per_lane;
  B->order = order;
  B->xcd = xcd != 0;
  if (wave_entity < 0 || wave_entity > 2)
    MPAS_FAIL(MPAS_EINVAL, "mpas_batch_config: wave_entity is 0, 1 or 2");
  B->wave_entity = wave_entity;
  return build_tables(B);
}

int mpas_batch_get_config(const mpas_batch *B, int32_t *block, int *members_per_lane, int *order,
                          int *xcd, int *wave_entity) {
  if (!B) MPAS_FAIL(MPAS_EINVAL, "batch is NULL");
  if (block) std::copy(B->block, B->block + kBatchKernels, block);
  if (members_per_lane) *members_per_lane = B->lpl;
  if (order) *order = B->order;
  if (xcd) *xcd = B->xcd;
  if (wave_entity) *wave_entity = B->wave_entity;
  return 0;
}

int mpas_batch_set_order(mpas_batch *B, int order) {
  if (int rc = check_batch(B)) return rc;
  if (order < MPAS_ORDER_GIVEN || order > MPAS_ORDER_MORTON)
    MPAS_FAIL(MPAS_EINVAL, "mpas_batch_set_order: unknown order");
  if (order == B->mesh_order) return 0;
  MPAS_HIP(hipStreamSynchronize(B->stream));
  void *save = nullptr;
  if (hipMalloc(&save, B->arena_bytes) != hipSuccess)
    MPAS_FAIL(MPAS_ENOMEM, "mpas_batch_set_order: no memory for the state copy");
  int rc = 0;
  if (hipMemcpyAsync(save, B->arena, B->arena_bytes, hipMemcpyDeviceToDevice, B->stream) !=
      hipSuccess) {
    set_error("mpas_batch_set_order: state copy failed");
    rc = MPAS_EHIP;
  }
  if (!rc) rc = permute_state(B, static_cast<const char *>(save), B->mesh_order, order);
  (void)hipStreamSynchronize(B->stream);
  (void)hipFree(save);
  return rc;
}

int mpas_batch_get_order(const mpas_batch *B) {
  if (!B) MPAS_FAIL(MPAS_EINVAL, "batch is NULL");
  return B->mesh_order;
}

int mpas_batch_tuned_mode(const mpas_batch *B) {
  if (!B) MPAS_FAIL(MPAS_EINVAL, "batch is NULL");
  return B->tuned_mode;
}

int mpas_batch_get_layout_blocks(const mpas_batch *B, int32_t *blocks) {
  if (!B || !blocks) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_get_layout_blocks: NULL argument");
  static_assert(MPAS_BATCH_KERNELS == kBatchKernels && MPAS_BATCH_LAYOUTS == kBatchModes,
                "ABI sizes of the per-layout table");
  for (int k = 0; k < kBatchKernels; k++)
    for (int m = 0; m < kBatchModes; m++) blocks[k * kBatchModes + m] = B->mblock[k][m];
  return 0;
}

int mpas_batch_set_layout_blocks(mpas_batch *B, const int32_t *blocks) {
  if (int rc = check_batch(B)) return rc;
  if (!blocks) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_set_layout_blocks: blocks is NULL");
  for (int i = 0; i < kBatchKernels * kBatchModes; i++)
    if (blocks[i] && (blocks[i] < 64 || blocks[i] > kBatchMaxBlock || blocks[i] % 64))
      MPAS_FAIL(MPAS_EINVAL,
                "mpas_batch_set_layout_blocks: sizes are 0 or multiples of 64 up to 512");
  MPAS_HIP(hipStreamSynchronize(B->stream));
  for (int k = 0; k < kBatchKernels; k++)
    for (int m = 0; m < kBatchModes; m++) B->mblock[k][m] = blocks[k * kBatchModes + m];
  return build_tables(B);
}

int mpas_batch_autotune(mpas_batch *B, int trial_steps, char *report, size_t cap) {
  if (int rc = check_batch(B)) return rc;
  if (trial_steps < 1) MPAS_FAIL(MPAS_EINVAL, "mpas_batch_autotune: trial_steps < 1");
  for (int m = 0; m < B->nmem; m++)
    if (!B->ready[m]) MPAS_FAIL(MPAS_ESTATE, "member " + std::to_string(m) + " has no state");
  MPAS_HIP(hipStreamSynchronize(B->stream));
  // start from the kernels' own sizes: per-layout sizes of an earlier run
  // would otherwise hide the per-kernel trials below
  for (auto &row : B->mblock)
    for (int &b : row) b = 0;
  // the trials advance the members: keep the state and restore it after
  void *save = nullptr;
  if (hipMalloc(&save, B->arena_bytes) != hipSuccess)
    MPAS_FAIL(MPAS_ENOMEM, "mpas_batch_autotune: no memory for the state copy");
  MPAS_HIP(hipMemcpyAsync(save, B->arena, B->arena_bytes, hipMemcpyDeviceToDevice, B->stream));
  std::string log = "[";
  int rc = 0;
  auto trial = [&](const char *what) -> double {
    if (rc) return 1e30;
    if ((rc = build_tables(B))) return 1e30;
    const double t = time_steps(B, trial_steps, &rc);
    char line[320];
    std::snprintf(line, sizeof line,
                  "%s{\"trial\": \"%s\", \"block\": [%d, %d, %d, %d, %d, %d], "
                  "\"members_per_lane\": %d, \"wave_entity\": %d, \"order\": %d, \"xcd\": %d, "
                  "\"mesh_order\": %d, \"ms_per_step\": %.4f, \"layout_block\": [",
                  log.size() > 1 ? ", " : "", what, B->block[0], B->block[1], B->block[2],
                  B->block[3], B->block[4], B->block[5], B->lpl, B->wave_entity, B->order,
                  B->xcd, B->mesh_order, 1e3 * t / trial_steps);
    log += line;
    for (int k = 0; k < kBatchKernels; k++)
      for (int m = 0; m < kBatchModes; m++)
        log += std::to_string(B->mblock[k][m]) +
               (k == kBatchKernels - 1 && m == kBatchModes - 1 ? "]}" : ", ");
    return rc ? 1e30 : t;
  };
  double best = trial("start");
  // the mesh order first, each candidate from the saved state (the start
  // trial measured the starting order: every other order is tried once)
  const int start_order = B->mesh_order;
  for (int o = MPAS_ORDER_GIVEN; o <= MPAS_ORDER_MORTON && !rc; o++) {
    if (o == start_order) continue;
    const int keep = B->mesh_order;
    if ((rc = permute_state(B, static_cast<const char *>(save), start_order, o))) break;
    const double t = trial("mesh order");
    if (t < best) best = t;
    else rc = permute_state(B, static_cast<const char *>(save), start_order, keep);
  }
  size_t gmax = 1;
  for (const GroupInfo &G : B->groups) gmax = std::max(gmax, G.members.size());
  // entity per wave and members per lane, placement and order, then each
  // kernel's workgroup size
  for (int we = 0; we < 3; we++)
    for (int lpl : {1, 2, 4, 8, 16}) {
      if ((lpl == B->lpl && we == B->wave_entity) || size_t(lpl) > gmax) continue;
      const int keep = B->lpl, kwe = B->wave_entity;
      B->lpl = lpl;
      B->wave_entity = we;
      const double t = trial("members_per_lane");
      if (t < best) best = t;
      else B->lpl = keep, B->wave_entity = kwe;
    }
  for (int x = 0; x < 2; x++)
    for (int o = 0; o < 2; o++) {
      if (x == B->xcd && o == B->order) continue;
      const int px = B->xcd, po = B->order;
      B->xcd = x;
      B->order = o;
      const double t = trial("placement");
      if (t < best) best = t;
      else B->xcd = px, B->order = po;
    }
  const int sizes[] = {64, 128, 256, 512};
  for (int k = 0; k < kBatchKernels; k++) {
    if (kernel_items(B, k, 0) == 0) continue;
    const int keep = B->block[k];
    int pick = keep;
    for (int bs : sizes) {
      if (bs == keep) continue;
      B->block[k] = bs;
      const double t = trial("block");
      if (t < best) best = t, pick = bs;
    }
    B->block[k] = pick;
  }
  // then per layout, where a kernel runs in more than one
  for (int k = 0; k < kBatchKernels; k++) {
    int present = 0;
    for (int m = 0; m < kBatchModes; m++) present += B->grid[k][m] > 0;
    if (present < 2) continue;
    for (int m = 0; m < kBatchModes; m++) {
      if (!B->grid[k][m]) continue;
      int pick = 0;
      for (int bs : sizes) {
        if (bs == B->block[k]) continue;
        B->mblock[k][m] = bs;
        const double t = trial("layout block");
        if (t < best) best = t, pick = bs;
      }
      B->mblock[k][m] = pick;
    }
  }
  // then the launch: the graph of per-phase launches or one persistent launch
  B->tuned_mode = MPAS_BATCH_GRAPH;
  if (!rc && persistent_supported(B)) {
    if (!(rc = build_tables(B))) {
      const double t = time_steps(B, trial_steps, &rc, MPAS_BATCH_PERSISTENT);
      char line[160];
      std::snprintf(line, sizeof line,
                    ", {\"trial\": \"persistent launch\", \"ms_per_step\": %.4f, "
                    "\"graph_ms_per_step\": %.4f}",
                    1e3 * t / trial_steps, 1e3 * best / trial_steps);
      log += line;
      if (!rc && t < best) B->tuned_mode = MPAS_BATCH_PERSISTENT;
    }
  }
  if (!rc) rc = build_tables(B);
  if (B->mesh_order == start_order) {
    (void)hipMemcpyAsync(B->arena, save, B->arena_bytes, hipMemcpyDeviceToDevice, B->stream);
  } else if (!rc) {
    rc = permute_state(B, static_cast<const char *>(save), start_order, B->mesh_order);
  }
  (void)hipStreamSynchronize(B->stream);
  (void)hipFree(save);
  if (rc) return rc;
  log += "]";
  if (report && cap) {
    if (log.size() + 1 > cap) MPAS_FAIL(MPAS_ESHAPE, "mpas_batch_autotune: report buffer too small");
    std::memcpy(report, log.c_str(), log.size() + 1);
  }
  return 0;
}

}  // extern "C"
