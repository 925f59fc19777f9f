/*
 * mpas_io.h — MPAS mesh files (HDF5 / netCDF-4 layout) for the MI355X
 * library: read a mesh into an mpas_mesh_desc, write one.
 *
 * The reference opens its mesh with HDF5 and reads each field as a dataset
 * by its MPAS name (mpasoflecsi/tasks/io.cc:14-158, io/helpers.hh:160-184;
 * the connectivity in mpasoflecsi/specialization/definition.hh:1080-1105).
 * The reader keeps the file's conventions (1-based ids, 0 = padding, rows of
 * the file's stride): mpas_ctx_create then drops the zeros and rebases the
 * ids as read_connectivity_2D does (definition.hh:670-718).  Only HDF5-based
 * files (netCDF-4) are read; a netCDF-3 classic file needs converting first
 * (nccopy -k nc4).  Status codes and mpas_last_error-like messages as in
 * mpas_hip.h; this library has no HIP dependency.
 */
#ifndef MPAS_IO_H
#define MPAS_IO_H

#include <stdint.h>

#include "mpas_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mpas_mesh_file mpas_mesh_file;

const char *mpas_io_last_error(void);

/* Reads every field mpas_ctx_create uses (connectivity, areas, lengths,
   kite areas, weights, positions; fVertex, meshDensity optional) and keeps
   them until mpas_mesh_file_close. */
int mpas_mesh_file_open(const char *path, mpas_mesh_file **out);
int mpas_mesh_file_close(mpas_mesh_file *f);
/* nCells, nEdges, nVertices, maxEdges, maxEdges2, vertexDegree */
int mpas_mesh_file_dims(const mpas_mesh_file *f, int64_t dims[6]);
/* The global attribute sphere_radius (0 when absent). */
int mpas_mesh_file_radius(const mpas_mesh_file *f, double *radius);
/* The mesh as mpas_ctx_create takes it; the pointers stay valid until
   mpas_mesh_file_close. */
int mpas_mesh_file_desc(const mpas_mesh_file *f, mpas_mesh_desc *d);
/* Shape of any dataset (rows, columns; columns = 1 for a 1-D dataset), or
   MPAS_EINVAL when the file has no such dataset. */
int mpas_mesh_file_shape(const mpas_mesh_file *f, const char *name, int64_t shape[2]);
/* Any dataset, converted to int32 (is_int) or double; n = rows * columns. */
int mpas_mesh_file_read(const mpas_mesh_file *f, const char *name, void *out, int64_t n,
                        int is_int);

/* One dataset of a file to write (row-major rows x columns, columns = 1 for
   a 1-D dataset). */
typedef struct mpas_io_dataset {
  const char *name;
  const void *data;
  int64_t rows, cols;
  int32_t is_int;
} mpas_io_dataset;
/* Writes an MPAS mesh file: the datasets and the global attributes
   sphere_radius and on_a_sphere = "YES". */
int mpas_mesh_file_write(const char *path, double sphere_radius, int n,
                         const mpas_io_dataset *datasets);

/* ---- output time series ----------------------------------------------------
   The run's output file, as apps/sw/run.cc:73-113 writes it.  The caller
   copies the fields to the host only at output steps. */
typedef struct mpas_output mpas_output;

/* What mesh_output_init (mpasoflecsi/tasks/io.cc:160-244) writes once. */
typedef struct mpas_output_mesh {
  int64_t nCells, nEdges, nVertices;
  int64_t vertexDegree, nVertLevels, nTracers;
  const double *areaCell, *xCell, *yCell, *zCell, *latCell, *lonCell; /* [nCells]    */
  const double *latVertex, *lonVertex;                                /* [nVertices] */
  const double *meshDensity;   /* [nCells], written as "mDensity"; NULL = 1        */
  const double *dvEdge, *dcEdge;                                      /* [nEdges]    */
  const int32_t *cellsOnVertex; /* [nVertices * vertexDegree], 1-based             */
} mpas_output_mesh;

/* mesh_output_init: creates the file with the attribute on_a_sphere = "YES",
   the dimension scales nCells, nEdges, nVertices, vertexDegree, nVertLevels,
   Time (num_output_times long) and nTracers, and the mesh fields above. */
int mpas_output_create(const char *path, int num_output_times,
                       const mpas_output_mesh *mesh, mpas_output **out);

enum { MPAS_OUT_CELLS = 0, MPAS_OUT_EDGES = 1, MPAS_OUT_VERTICES = 2 };
/* One field of one output time (field_output_wrapper::output,
   mpasoflecsi/tasks/io.hh:13-47): data[entity][level][tracer] with
   levels = 0 for a field without levels (h_s) and tracers = 0 for one
   without tracers. */
typedef struct mpas_output_field {
  const char *name; /* h, h_s, vorticity, potential_vorticity, tracers */
  const double *data;
  int32_t space;    /* MPAS_OUT_CELLS / _EDGES / _VERTICES */
  int32_t levels, tracers;
} mpas_output_field;
/* Writes time row t of the field's dataset [Time][entity](level)(tracer)
   (io/helpers.hh:127-258); the dataset is created, with its dimension
   scales attached, at the field's first write.  t must lie in
   [0, num_output_times). */
int mpas_output_write(mpas_output *o, int t, const mpas_output_field *field);
/* mesh_output_finalize (tasks/io.cc:247-251). */
int mpas_output_close(mpas_output *o);

/* Reading an output file back (post-processing, tests): the rank and
   dimensions of a dataset, and, when out is not NULL, all its n entries
   converted to double. */
int mpas_output_read(const char *path, const char *name, int *rank, int64_t dims[4],
                     double *out, int64_t n);
/* Name of the dimension scale attached to dimension dim of a dataset
   (empty when none is). */
int mpas_output_scale_name(const char *path, const char *name, int dim, char *buf,
                           int64_t len);

#ifdef __cplusplus
}
#endif
#endif /* MPAS_IO_H */
