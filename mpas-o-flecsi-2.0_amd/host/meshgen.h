// Icosahedral spherical Voronoi (TRiSK) mesh generator.
//
// Produces every MPAS mesh field that the shallow-water / FCT path reads
// (mpasoflecsi/tasks/io.cc:14-158 lists them), in MPAS file conventions:
// connectivity is 1-based with 0 as padding, row strides maxEdges (=8) and
// maxEdges2 (=16) as in mpasoflecsi/common/constants.hh:11-12.
//
// The generator is needed because the reference's meshes
// (data/meshes/Sphere_240QU.nc, Sphere_480QU.nc) are not available.
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mg_mesh mg_mesh;

enum mg_order { MG_ORDER_NATIVE = 0, MG_ORDER_HILBERT = 1 };

// level: icosahedral subdivision level (cells = 10*4^level + 2).
// radius: sphere radius in metres.
// lloyd_iters: Lloyd (centroidal) iterations applied to the generators.
// order: numbering of cells/edges/vertices in the produced "file".
// omega: rotation rate used for fVertex/fEdge/fCell = 2*omega*sin(lat).
// Returns NULL on failure (mg_last_error() says why).
mg_mesh *mg_generate(int level, double radius, int lloyd_iters, int order,
                     double omega);
// rotate = 0 keeps the icosahedron with a vertex at each pole (the usual
// orientation of MPAS quasi-uniform meshes); mg_generate uses rotate = 1,
// a fixed generic rotation that keeps arc midpoints off the coordinate
// planes the reference's FCT setup may not meet.
mg_mesh *mg_generate_oriented(int level, double radius, int lloyd_iters, int order,
                              double omega, int rotate);
void mg_free(mg_mesh *m);
const char *mg_last_error(void);

// Dimensions: "nCells", "nEdges", "nVertices", "maxEdges", "maxEdges2",
// "vertexDegree".
int64_t mg_dim(const mg_mesh *m, const char *name);
// Integer fields (int32): nEdgesOnCell, edgesOnCell, cellsOnCell,
// verticesOnCell, cellsOnEdge, verticesOnEdge, nEdgesOnEdge, edgesOnEdge,
// cellsOnVertex, edgesOnVertex.  Sets *n to the element count.
const int32_t *mg_int_field(const mg_mesh *m, const char *name, int64_t *n);
// Double fields: xCell yCell zCell latCell lonCell xEdge yEdge zEdge latEdge
// lonEdge xVertex yVertex zVertex latVertex lonVertex areaCell areaTriangle
// kiteAreasOnVertex dcEdge dvEdge angleEdge weightsOnEdge meshDensity
// fVertex fEdge fCell.
const double *mg_dbl_field(const mg_mesh *m, const char *name, int64_t *n);

#ifdef __cplusplus
}
#endif
