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
  // edges
  const int2 *e2c;  // m.cells(e): ascending cell id (transpose_cnx)
  const int2 *e2v;  // m.vertices(e): file order
  const double *dc, *dv;
  const uint8_t *nEoE;
  const int *eoe;     // [j][E] m.edges(e), file order
  const double *woe;  // [j][E] weightsOnEdge
  // cells
  const double *area;
  const uint8_t *nEoC;
  const int *c2e;        // [j][C] m.edges(c), file order
  const int *c2c;        // [j][C] m.cells(c), file order
  const int *cOpp;       // [j][C] the other cell of edge c2e[j]
  const uint8_t *sgnNeg;    // bit j: c == m.cells(c2e[j])[0], edgeSignOnCell = -1
  const uint32_t *ascPerm;  // nibble r: slot j of the r-th smallest edge id
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
  const int *mlc;     // maxLevelCell [C]; nullptr = nL everywhere
  const int *mlet;    // maxLevelEdgeTop [E]; nullptr = nL everywhere
  const int *hoMask;  // highOrderAdvMask [E][nL]; nullptr = 1 everywhere
  double msd;         // meshScalingDel2: one value for every edge
  double gravity, coef3rd;
};

// std::max / std::min exactly as <algorithm> defines them (the result for
// signed zeros and NaN follows the comparison, unlike fmax / fmin).
__device__ __forceinline__ double smax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double smin(double a, double b) { return (b < a) ? b : a; }

__device__ __forceinline__ int lvl_cell(const DMesh &M, int c) {
  return M.mlc ? M.mlc[c] : M.nL;
}
__device__ __forceinline__ int lvl_edge(const DMesh &M, int e) {
  return M.mlet ? M.mlet[e] : M.nL;
}
__device__ __forceinline__ double sign_on_cell(uint8_t neg, int j) {
  return ((neg >> j) & 1) ? -1.0 : 1.0;
}
__device__ __forceinline__ int v2e_id(int x) { return x & 0x7fffffff; }
__device__ __forceinline__ double v2e_sign(int x) { return x < 0 ? -1.0 : 1.0; }

// Grid-stride helpers: one work item per lane.
__device__ __forceinline__ int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

}  // namespace mpasx
