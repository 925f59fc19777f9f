// Device view of the flattened mesh and the arithmetic helpers shared by
// the per-task and the stage-fused kernels.
//
// Layout in HBM (all int32 ids, FP64 values):
//  * per-entity scalars: plain arrays;
//  * ragged lists: slot-major ELL, list[j * n + i], so lane i of a wavefront
//    reads consecutive words for every slot j (one coalesced stream per slot);
//  * fields: the reference's layouts, vltensor [n][nLevels] and vltracer
//    [nCells][nLevels][nTracers] (apps/sw/state.hh:75-87), so a gathered
//    cell brings all of its tracers in one 8*nTracers-byte run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpasx {

struct DMesh {
  int nC, nE, nV, nL, nT;
  int wC;    // ELL width of the per-cell lists (max edges of a cell, <= 8)
  int wEoE;  // ELL width of edgesOnEdge
  int wAdv;  // ELL width of advCellsForEdge
  // Padded widths the fused kernels read unconditionally: 6 / 10 when every
  // cell has <= 6 edges and every stencil <= 10 entries, else 8 / 16.  The
  // ELL arrays are allocated with (at least) these widths.
  int narrow;
  // edges
  const int2 *e2c;  // m.cells(e): ascending cell id (transpose_cnx)
  const int2 *e2v;  // m.vertices(e): file order
  const double *dc, *dv;
  const uint8_t *nEoE;
  const int *eoe;     // [j][E] m.edges(e), file order
  const double *woe;  // [j][E] weightsOnEdge
  // the same two slots at a time, [j/2][E][2] (the one-level kernels read a
  // slot pair with one access)
  const int *eoeP;
  const double *woeP;
  // cells
  const double *area;
  const uint8_t *nEoC;
  const int *c2e;        // [j][C] m.edges(c), file order
  const int *c2c;        // [j][C] m.cells(c), file order
  const int *cOpp;       // [j][C] the other cell of edge c2e[j]
  const uint8_t *sgnNeg;    // bit j: c == m.cells(c2e[j])[0], edgeSignOnCell = -1
  const uint32_t *ascPerm;  // nibble r: slot j of the r-th smallest edge id
  const uint32_t *ascRank;  // nibble j: rank of slot j's edge id (15 past the count)
  // vertices
  const double *areaTri, *fV;
  const double *kite;  // [i][V] kiteAreasOnVertex, file order
  const int *v2c;      // [i][V] m.cells(v): ascending
  const int *v2e;      // [i][V] m.edges(v): ascending; bit 31 set when
                       // v == m.vertices(e)[0] (circulation sign -1)
  // FCT / del2 (init_functions.cc products)
  const uint8_t *nAdv;
  const int *adv;                       // [j][E] advCellsForEdge
  const double *advCoef, *advCoef3;     // [j][E] advCoef, advCoef_3rd
  const int *advP;                      // advCellsForEdge, [j/2][E][2]
  int advC12;  // 1: every edge's advection cells start with cells(e)[0], [1]
  const double *advCoefP, *advCoef3P;   // advCoef, advCoef_3rd, [j/2][E][2]
  // advCoef_3rd in compact form (one-level kernels): per edge, 2 bits per
  // slot say whether the slot's advCoef_3rd is advCoef (1), -advCoef (2) -
  // true bit for bit of every slot only one of the edge's two cell fits
  // holds - or the next of the edge's explicit values (0), which are
  // [kAdvExplicit/2][E][2]; nullptr when some edge has more explicit slots
  const uint32_t *advSide;
  const double *advC3xP;
  // LDS staging of the one-level cells kernel: the cells of logical block
  // b (kCellsBlock lanes, stgCpb cells) touch the edges stgEdge[stgOff[b]
  // .. stgOff[b + 1]) (ascending); the 10-bit field j of c2eLoc[c] is the
  // position of edge slot j of cell c in its block's list.  nullptr: none.
  const uint64_t *c2eLoc;
  const int *stgOff, *stgEdge;
  int stgMax, stgCpb;  // the largest block list; cells per block
  // LDS staging of the one-level momentum kernel: the edges of logical
  // block b (kEdgesBlock lanes) gather their edgesOnEdge values from the
  // distinct edges eStgEdge[eStgOff[b] .. eStgOff[b + 1]) (ascending);
  // e2eLoc[2 e + j / 6] holds in bits 10 (j % 6) .. the position of slot j
  // of edge e in its block's list.  nullptr: none.
  const uint64_t *e2eLoc;
  const int *eStgOff, *eStgEdge;
  int eStgMax, eStgLen;  // the largest block list; all lists
  // LDS staging of the one-level high-order flux kernel, the same for the
  // advection cells: the aStgEpb edges of logical block b (kHoBlock lanes)
  // read the tracers of cells aStgCell[aStgOff[b] .. aStgOff[b + 1]);
  // a2cLoc[2 e + i / 6] holds the 10-bit position of advection slot i.
  const uint64_t *a2cLoc;
  const int *aStgOff, *aStgCell;
  int aStgMax, aStgLen, aStgEpb;
  // LDS staging of the one-level tracer update: the tStgCpb cells of
  // logical block b (kTracerBlock lanes, cells from b * tStgCpb) read the
  // limiter factors and tracers of the cells tStgCell[tStgOff[b] ..
  // tStgOff[b + 1]): the block's own cells first, in order, then the
  // other cells across their edges, ascending; the 10-bit field j of
  // c2cLoc[c] is the position of the cell across edge slot j of cell c
  // (slots past the cell's count: its own position).  nullptr: none.
  const uint64_t *c2cLoc;
  const int *tStgOff, *tStgCell;
  int tStgMax, tStgLen, tStgCpb;
  // The same lists for the one-level FCT cells kernel's workgroups
  // (kCellsBlock lanes, cStgCpb cells), which stage their cells' and their
  // neighbours' tracers and thickness.
  const uint64_t *c2cLocC;
  const int *cStgOff, *cStgCell;
  int cStgMax, cStgLen, cStgCpb;
  const int *mlc;     // maxLevelCell [C]; nullptr = nL everywhere
  const int *mlet;    // maxLevelEdgeTop [E]; nullptr = nL everywhere
  const int *hoMask;  // highOrderAdvMask [E][nL]; nullptr = 1 everywhere
  double msd;         // meshScalingDel2: one value for every edge
  double gravity, coef3rd;
};

// Workgroup size of the one-level cells kernel (its LDS staging groups the
// cells by workgroup).
constexpr int kCellsBlock = 256;
// Workgroup size of the one-level momentum kernel (its LDS staging groups
// the edges by workgroup).
constexpr int kEdgesBlock = 256;
// The same for the high-order flux kernel, and its tracers per lane (four
// where the count allows: each lane streams its edge's stencil once for
// all of them; else pairs when the count is even).
constexpr int kHoBlock = 256;
#ifndef MPAS_TPL_HO
#define MPAS_TPL_HO 4
#endif
inline int ho_lanes_per_edge(int nT) {
  return nT % MPAS_TPL_HO == 0 ? nT / MPAS_TPL_HO : nT % 2 == 0 ? nT / 2 : nT;
}

// The same for the tracer update, whose lanes take MPAS_TPL_TRACER
// tracers each (one: four lanes per cell at four tracers).
constexpr int kTracerBlock = 256;
#ifndef MPAS_TPL_TRACER
#define MPAS_TPL_TRACER 1
#endif
inline int tracer_lanes_per_cell(int nT) {
  return nT % MPAS_TPL_TRACER == 0 ? nT / MPAS_TPL_TRACER : nT % 2 == 0 ? nT / 2 : nT;
}

// explicit advCoef_3rd values per edge in the compact form: the two cells of
// the edge and their two common neighbours (the slots both cell fits hold)
constexpr int kAdvExplicit = 4;

// std::max / std::min exactly as <algorithm> defines them (the result for
// signed zeros and NaN follows the comparison, unlike fmax / fmin).
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }


__device__ __forceinline__ double sign_on_cell(uint8_t neg, int j) {
  return ((neg >> j) & 1) ? -1.0 : 1.0;
}
__device__ __forceinline__ int v2e_id(int x) { return x & 0x7fffffff; }
__device__ __forceinline__ double v2e_sign(int x) { return x < 0 ? -1.0 : 1.0; }

// Loads and stores at a 32-bit element index from a base held in scalar
// registers: the byte offset is formed in 32 bits, which lets the compiler
// use the scalar-base form of global memory instructions (one VGPR per
// address instead of two).  Every array handed to the kernels is below
// 4 GiB (checked when the context is created).  The access is made in the
// global address space explicitly: pointers that reach a kernel through a
// descriptor in memory (the ensemble kernels) are otherwise generic, and
// generic (flat) accesses are slower and wait on LDS traffic as well.
// (The host compilation pass only parses these bodies; it keeps the
// generic form, which its vector types require.)
#ifdef __HIP_DEVICE_COMPILE__
#define MPAS_GLOBAL __attribute__((address_space(1)))
#else
#define MPAS_GLOBAL
#endif
template <class T>
__device__ __forceinline__ T ld(const T *__restrict__ p, uint32_t i) {
  const MPAS_GLOBAL char *b = (const MPAS_GLOBAL char *)(p);
  return *(const MPAS_GLOBAL T *)(b + i * uint32_t(sizeof(T)));
}
// Loads of the static mesh arrays (never written while a kernel runs) in
// the constant address space: when the index is the same for every lane of
// a wave (one mesh entity per wave, as the ensemble kernels arrange it) the
// compiler turns them into scalar loads, which leave the vector memory path
// to the state; with a per-lane index they are ordinary vector loads.
#ifdef __HIP_DEVICE_COMPILE__
#define MPAS_CONST __attribute__((address_space(4)))
#else
#define MPAS_CONST
#endif
template <class T>
__device__ __forceinline__ T ldc(const T *__restrict__ p, uint32_t i) {
  const MPAS_CONST char *b = (const MPAS_CONST char *)(p);
  return *(const MPAS_CONST T *)(b + i * uint32_t(sizeof(T)));
}
// Record i of V-sized records (a native 2-vector of S) the same way.
template <class V, class S>
__device__ __forceinline__ V ldc_vec(const S *__restrict__ p, uint32_t i) {
  const MPAS_CONST char *b = (const MPAS_CONST char *)(p);
  return *(const MPAS_CONST V *)(b + i * uint32_t(sizeof(V)));
}
// Streams read or written once per launch (the slot-major stencils, a
// lane's own state entries): with NT true, non-temporal accesses, which
// spare the caches for the neighbour gathers.  The switches select which
// streams of the one-entity-per-lane launches take them.
#ifndef MPAS_NT_STENCIL
#define MPAS_NT_STENCIL 1
#endif
#ifndef MPAS_NT_CELLS
#define MPAS_NT_CELLS 1
#endif
#ifndef MPAS_NT_NEXT
#define MPAS_NT_NEXT 1
#endif
constexpr bool kNtStencil = MPAS_NT_STENCIL, kNtCells = MPAS_NT_CELLS, kNtNext = MPAS_NT_NEXT;
template <bool NT, class T>
__device__ __forceinline__ T ldn(const T *__restrict__ p, uint32_t i) {
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (NT) {
    const MPAS_GLOBAL char *b = (const MPAS_GLOBAL char *)(p);
    return __builtin_nontemporal_load((const MPAS_GLOBAL T *)(b + i * uint32_t(sizeof(T))));
  }
#endif
  return ldc(p, i);
}
// Record i of V-sized records (V a native 2-vector of the element type),
// non-temporal with NT: one access for two consecutive slots of a
// slot-paired row.
typedef int mpas_i2v __attribute__((ext_vector_type(2)));
typedef double mpas_d2v __attribute__((ext_vector_type(2)));
typedef unsigned long long mpas_u2v __attribute__((ext_vector_type(2)));
template <bool NT, class V, class S>
__device__ __forceinline__ V ld_vec(const S *__restrict__ p, uint32_t i) {
  const MPAS_GLOBAL char *b = (const MPAS_GLOBAL char *)(p);
  const MPAS_GLOBAL V *q = (const MPAS_GLOBAL V *)(b + i * uint32_t(sizeof(V)));
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (NT) return __builtin_nontemporal_load(q);
#endif
  return *q;
}
template <bool NT, class T>
__device__ __forceinline__ T ldo(const T *__restrict__ p, uint32_t i) {
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (NT) {
    const MPAS_GLOBAL char *b = (const MPAS_GLOBAL char *)(p);
    return __builtin_nontemporal_load((const MPAS_GLOBAL T *)(b + i * uint32_t(sizeof(T))));
  }
#endif
  return ld(p, i);
}
template <class T>
__device__ __forceinline__ void st(T *p, uint32_t i, T v) {
  MPAS_GLOBAL char *b = (MPAS_GLOBAL char *)(p);
  *(MPAS_GLOBAL T *)(b + i * uint32_t(sizeof(T))) = v;
}
template <bool NT, class T>
__device__ __forceinline__ void stn(T *p, uint32_t i, T v) {
#ifdef __HIP_DEVICE_COMPILE__
  if constexpr (NT) {
    MPAS_GLOBAL char *b = (MPAS_GLOBAL char *)(p);
    __builtin_nontemporal_store(v, (MPAS_GLOBAL T *)(b + i * uint32_t(sizeof(T))));
    return;
  }
#endif
  st(p, i, v);
}

// TPL consecutive tracers of one entity: one 16-byte load / store per pair
// (the offset i is even whenever TPL is: NT is a multiple of TPL then).
template <int TPL>
__device__ __forceinline__ void ld_group(const double *__restrict__ p, uint32_t i,
                                         double (&v)[TPL]) {
  if constexpr (TPL % 2 == 0) {
#pragma unroll
    for (int g = 0; g < TPL; g += 2) {
      const double2 x = ld(reinterpret_cast<const double2 *>(p), (i + g) / 2);
      v[g] = x.x;
      v[g + 1] = x.y;
    }
  } else {
#pragma unroll
    for (int g = 0; g < TPL; g++) v[g] = ld(p, i + g);
  }
}
template <int TPL>
__device__ __forceinline__ void st_group(double *p, uint32_t i, const double (&v)[TPL]) {
  if constexpr (TPL % 2 == 0) {
#pragma unroll
    for (int g = 0; g < TPL; g += 2)
      st(reinterpret_cast<double2 *>(p), (i + g) / 2, make_double2(v[g], v[g + 1]));
  } else {
#pragma unroll
    for (int g = 0; g < TPL; g++) st(p, i + g, v[g]);
  }
}

// Level pairs per lane where an entity's levels (or ensemble members) lie
// on adjacent lanes and their count is even: the momentum, high-order flux
// and diagnostic kernels take two levels per lane, one 16-byte access per
// pair.
#ifndef MPAS_LEVEL_PAIRS
#define MPAS_LEVEL_PAIRS 1
#endif

// LP consecutive levels of one entity from p + i: one 16-byte load / store
// per level pair (i is even: the level count is, and a lane's first level).
template <int LP>
__device__ __forceinline__ void ld_lv(const double *__restrict__ p, uint32_t i,
                                      double (&v)[LP]) {
  ld_group<LP>(p, i, v);
}
template <int LP>
__device__ __forceinline__ void st_lv(double *p, uint32_t i, const double (&v)[LP]) {
  st_group<LP>(p, i, v);
}

__device__ __forceinline__ int lvl_cell(const DMesh &M, int c) {
  return M.mlc ? ldc(M.mlc, uint32_t(c)) : M.nL;
}
__device__ __forceinline__ int lvl_edge(const DMesh &M, int e) {
  return M.mlet ? ldc(M.mlet, uint32_t(e)) : M.nL;
}

// Grid-stride helpers: one work item per lane.
__device__ __forceinline__ int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Work item of this lane under XCD-aware placement.  The MI355X hands
// workgroups to its 8 XCDs round-robin (workgroup b runs on XCD b % 8, each
// XCD with its own L2).  Giving XCD x the contiguous range of logical blocks
// [x * per, (x + 1) * per) keeps neighbouring mesh entities (consecutive ids
// along the Hilbert numbering) in one L2, so data shared across block
// boundaries is fetched once instead of once per XCD.  The grid must be
// 8 * per blocks (xcd_grid); items past n return.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_item() {
  const int per = gridDim.x / kXcds;
  const int b = blockIdx.x;
  return ((b % kXcds) * per + b / kXcds) * blockDim.x + threadIdx.x;
}
// The same for block b of a range of nb blocks (nb a multiple of the XCD
// count) within a launch whose earlier ranges also hold multiples of it:
// block b of the range runs on XCD b % 8 as well.
__device__ __forceinline__ int xcd_item_in(int b, int nb) {
  const int per = nb / kXcds;
  return ((b % kXcds) * per + b / kXcds) * blockDim.x + threadIdx.x;
}
inline int xcd_grid(int64_t n, int bs) {
  const int64_t blocks = (n + bs - 1) / bs;
  return int((blocks + kXcds - 1) / kXcds * kXcds);
}

}  // namespace mpasx
