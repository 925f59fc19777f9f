// MPAS mesh files through the HDF5 C API: the datasets the reference reads
// by name (mpasoflecsi/tasks/io.cc:14-158), in the file's own conventions.
#include <hdf5.h>
#include <hdf5_hl.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <type_traits>
#include <vector>

#include "mpas_io.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// Closes an HDF5 handle when it goes out of scope.
struct Handle {
  hid_t id;
  herr_t (*close)(hid_t);
  ~Handle() {
    if (id >= 0) close(id);
  }
};

// Quiet HDF5's own error stack printing: failures come back as status codes.
void quiet() { H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); }

bool has_dataset(hid_t file, const char *name) {
  return H5Lexists(file, name, H5P_DEFAULT) > 0;
}

int shape_of(hid_t file, const char *name, int64_t shape[2]) {
  if (!has_dataset(file, name)) return fail(MPAS_EINVAL, std::string("no dataset ") + name);
  Handle ds{H5Dopen2(file, name, H5P_DEFAULT), H5Dclose};
  if (ds.id < 0) return fail(MPAS_EINVAL, std::string("cannot open dataset ") + name);
  Handle sp{H5Dget_space(ds.id), H5Sclose};
  const int rank = H5Sget_simple_extent_ndims(sp.id);
  if (rank < 1 || rank > 2)
    return fail(MPAS_ESHAPE, std::string(name) + " has rank " + std::to_string(rank));
  hsize_t dims[2] = {0, 1};
  H5Sget_simple_extent_dims(sp.id, dims, nullptr);
  shape[0] = int64_t(dims[0]);
  shape[1] = rank == 2 ? int64_t(dims[1]) : 1;
  return 0;
}

template <class T>
int read_all(hid_t file, const char *name, std::vector<T> &out, int64_t shape[2]) {
  if (int rc = shape_of(file, name, shape)) return rc;
  out.resize(size_t(shape[0] * shape[1]));
  Handle ds{H5Dopen2(file, name, H5P_DEFAULT), H5Dclose};
  const hid_t mem = std::is_same<T, int32_t>::value ? H5T_NATIVE_INT32 : H5T_NATIVE_DOUBLE;
  if (H5Dread(ds.id, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
    return fail(MPAS_EINVAL, std::string("cannot read dataset ") + name);
  return 0;
}

}  // namespace

struct mpas_mesh_file {
  hid_t file = -1;
  int64_t dims[6] = {};
  double radius = 0;
  std::map<std::string, std::vector<int32_t>> ints;
  std::map<std::string, std::vector<double>> dbls;
  ~mpas_mesh_file() {
    if (file >= 0) H5Fclose(file);
  }
};

extern "C" {

const char *mpas_io_last_error(void) { return g_err.c_str(); }

int mpas_mesh_file_open(const char *path, mpas_mesh_file **out) {
  if (!path || !out) return fail(MPAS_EINVAL, "mpas_mesh_file_open: NULL argument");
  *out = nullptr;
  quiet();
  if (H5Fis_hdf5(path) <= 0)
    return fail(MPAS_EINVAL, std::string(path) +
                                 ": not an HDF5 (netCDF-4) file (convert with nccopy -k nc4)");
  auto f = std::make_unique<mpas_mesh_file>();
  f->file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f->file < 0) return fail(MPAS_EINVAL, std::string("cannot open ") + path);
  int64_t sh[2];
  // connectivity (definition.hh:1080-1105) and the fields of io.cc:14-63
  const char *ints[] = {"verticesOnCell", "edgesOnCell", "cellsOnCell", "verticesOnEdge",
                        "edgesOnEdge"};
  for (const char *n : ints)
    if (int rc = read_all(f->file, n, f->ints[n], sh)) return rc;
  const char *dbls[] = {"areaCell", "dvEdge", "dcEdge", "areaTriangle", "kiteAreasOnVertex",
                        "weightsOnEdge"};
  for (const char *n : dbls)
    if (int rc = read_all(f->file, n, f->dbls[n], sh)) return rc;
  const char *optional[] = {"fVertex", "meshDensity", "xCell", "yCell", "zCell",
                            "xVertex", "yVertex", "zVertex"};
  for (const char *n : optional)
    if (has_dataset(f->file, n))
      if (int rc = read_all(f->file, n, f->dbls[n], sh)) return rc;
  int64_t a[2];
  shape_of(f->file, "edgesOnCell", a);
  f->dims[0] = a[0];
  f->dims[3] = a[1];
  shape_of(f->file, "edgesOnEdge", a);
  f->dims[1] = a[0];
  f->dims[4] = a[1];
  shape_of(f->file, "kiteAreasOnVertex", a);
  f->dims[2] = a[0];
  f->dims[5] = a[1];
  // shapes of the remaining fields against the dimensions
  auto expect = [&](const char *n, int64_t rows, int64_t cols) -> int {
    int64_t s[2];
    if (shape_of(f->file, n, s)) return 0;  // optional and absent
    if (s[0] != rows || s[1] != cols)
      return fail(MPAS_ESHAPE, std::string(n) + " has shape " + std::to_string(s[0]) + " x " +
                                   std::to_string(s[1]));
    return 0;
  };
  const int64_t C = f->dims[0], E = f->dims[1], V = f->dims[2];
  int rc = 0;
  if ((rc = expect("verticesOnCell", C, f->dims[3])) || (rc = expect("cellsOnCell", C, f->dims[3])) ||
      (rc = expect("verticesOnEdge", E, 2)) || (rc = expect("areaCell", C, 1)) ||
      (rc = expect("dvEdge", E, 1)) || (rc = expect("dcEdge", E, 1)) ||
      (rc = expect("areaTriangle", V, 1)) || (rc = expect("weightsOnEdge", E, f->dims[4])) ||
      (rc = expect("fVertex", V, 1)) || (rc = expect("meshDensity", C, 1)) ||
      (rc = expect("xCell", C, 1)) || (rc = expect("xVertex", V, 1)))
    return rc;
  if (H5Aexists(f->file, "sphere_radius") > 0) {
    Handle at{H5Aopen(f->file, "sphere_radius", H5P_DEFAULT), H5Aclose};
    H5Aread(at.id, H5T_NATIVE_DOUBLE, &f->radius);
  }
  *out = f.release();
  return 0;
}

int mpas_mesh_file_close(mpas_mesh_file *f) {
  delete f;
  return 0;
}

int mpas_mesh_file_dims(const mpas_mesh_file *f, int64_t dims[6]) {
  if (!f || !dims) return fail(MPAS_EINVAL, "mpas_mesh_file_dims: NULL argument");
  std::memcpy(dims, f->dims, sizeof(f->dims));
  return 0;
}

int mpas_mesh_file_radius(const mpas_mesh_file *f, double *radius) {
  if (!f || !radius) return fail(MPAS_EINVAL, "mpas_mesh_file_radius: NULL argument");
  *radius = f->radius;
  return 0;
}

int mpas_mesh_file_desc(const mpas_mesh_file *f, mpas_mesh_desc *d) {
  if (!f || !d) return fail(MPAS_EINVAL, "mpas_mesh_file_desc: NULL argument");
  std::memset(d, 0, sizeof(*d));
  d->nCells = int32_t(f->dims[0]);
  d->nEdges = int32_t(f->dims[1]);
  d->nVertices = int32_t(f->dims[2]);
  d->maxEdges = int32_t(f->dims[3]);
  d->maxEdges2 = int32_t(f->dims[4]);
  d->vertexDegree = int32_t(f->dims[5]);
  auto I = [&](const char *n) { return f->ints.at(n).data(); };
  auto D = [&](const char *n) -> const double * {
    auto it = f->dbls.find(n);
    return it == f->dbls.end() ? nullptr : it->second.data();
  };
  d->verticesOnCell = I("verticesOnCell");
  d->edgesOnCell = I("edgesOnCell");
  d->cellsOnCell = I("cellsOnCell");
  d->verticesOnEdge = I("verticesOnEdge");
  d->edgesOnEdge = I("edgesOnEdge");
  d->areaCell = D("areaCell");
  d->dvEdge = D("dvEdge");
  d->dcEdge = D("dcEdge");
  d->areaTriangle = D("areaTriangle");
  d->kiteAreasOnVertex = D("kiteAreasOnVertex");
  d->fVertex = D("fVertex");
  d->weightsOnEdge = D("weightsOnEdge");
  d->xCell = D("xCell");
  d->yCell = D("yCell");
  d->zCell = D("zCell");
  d->xVertex = D("xVertex");
  d->yVertex = D("yVertex");
  d->zVertex = D("zVertex");
  d->meshDensity = D("meshDensity");
  return 0;
}

int mpas_mesh_file_shape(const mpas_mesh_file *f, const char *name, int64_t shape[2]) {
  if (!f || !name || !shape) return fail(MPAS_EINVAL, "mpas_mesh_file_shape: NULL argument");
  return shape_of(f->file, name, shape);
}

int mpas_mesh_file_read(const mpas_mesh_file *f, const char *name, void *out, int64_t n,
                        int is_int) {
  if (!f || !name || !out) return fail(MPAS_EINVAL, "mpas_mesh_file_read: NULL argument");
  int64_t sh[2];
  if (int rc = shape_of(f->file, name, sh)) return rc;
  if (sh[0] * sh[1] != n)
    return fail(MPAS_ESHAPE, std::string(name) + " has " + std::to_string(sh[0] * sh[1]) +
                                 " entries, not " + std::to_string(n));
  Handle ds{H5Dopen2(f->file, name, H5P_DEFAULT), H5Dclose};
  if (H5Dread(ds.id, is_int ? H5T_NATIVE_INT32 : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL,
              H5P_DEFAULT, out) < 0)
    return fail(MPAS_EINVAL, std::string("cannot read dataset ") + name);
  return 0;
}

namespace {

int write_dataset(hid_t file, const mpas_io_dataset &x) {
  if (!x.name || !x.data || x.rows < 0 || x.cols < 1)
    return fail(MPAS_EINVAL, "bad dataset");
  const int rank = x.cols == 1 ? 1 : 2;
  const hsize_t dims[2] = {hsize_t(x.rows), hsize_t(x.cols)};
  Handle sp{H5Screate_simple(rank, dims, nullptr), H5Sclose};
  const hid_t ty = x.is_int ? H5T_STD_I32LE : H5T_IEEE_F64LE;
  Handle ds{H5Dcreate2(file, x.name, ty, sp.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
            H5Dclose};
  if (ds.id < 0) return fail(MPAS_EINVAL, std::string("cannot create dataset ") + x.name);
  if (H5Dwrite(ds.id, x.is_int ? H5T_NATIVE_INT32 : H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL,
               H5P_DEFAULT, x.data) < 0)
    return fail(MPAS_EINVAL, std::string("cannot write dataset ") + x.name);
  return 0;
}

void write_radius(hid_t file, double sphere_radius) {
  Handle sp{H5Screate(H5S_SCALAR), H5Sclose};
  Handle at{H5Acreate2(file, "sphere_radius", H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT),
            H5Aclose};
  H5Awrite(at.id, H5T_NATIVE_DOUBLE, &sphere_radius);
}

}  // namespace

int mpas_mesh_file_write(const char *path, double sphere_radius, int n,
                         const mpas_io_dataset *datasets) {
  if (!path || (n && !datasets)) return fail(MPAS_EINVAL, "mpas_mesh_file_write: NULL argument");
  quiet();
  Handle file{H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose};
  if (file.id < 0) return fail(MPAS_EINVAL, std::string("cannot create ") + path);
  for (int i = 0; i < n; i++)
    if (int rc = write_dataset(file.id, datasets[i])) return rc;
  write_radius(file.id, sphere_radius);
  {
    Handle sp{H5Screate(H5S_SCALAR), H5Sclose};
    Handle ty{H5Tcopy(H5T_C_S1), H5Tclose};
    H5Tset_size(ty.id, 3);
    Handle at{H5Acreate2(file.id, "on_a_sphere", ty.id, sp.id, H5P_DEFAULT, H5P_DEFAULT),
              H5Aclose};
    H5Awrite(at.id, ty.id, "YES");
  }
  return 0;
}

}  // extern "C"

struct mpas_output {
  hid_t file = -1;
  int ntimes = 0;
  int64_t n[3] = {0, 0, 0};  // cells, edges, vertices
  int64_t levels = 0, tracers = 0;
  ~mpas_output() {
    if (file >= 0) H5Fclose(file);
  }
};

namespace {

const char *const kSpace[3] = {"nCells", "nEdges", "nVertices"};

// A dimension scale: a zero-filled int dataset of the dimension's length
// marked as a scale (helper1d::write_dimscale, io/helpers.hh:82-118).
int dim_scale(hid_t file, const char *name, int64_t len) {
  std::vector<int32_t> zero(size_t(len > 0 ? len : 1), 0);  // a valid pointer when len = 0
  mpas_io_dataset d{name, zero.data(), len, 1, 1};
  if (int rc = write_dataset(file, d)) return rc;
  Handle ds{H5Dopen2(file, name, H5P_DEFAULT), H5Dclose};
  if (ds.id < 0 || H5DSset_scale(ds.id, name) < 0)
    return fail(MPAS_EINVAL, std::string("cannot make scale ") + name);
  return 0;
}

int attach(hid_t file, const char *dataset, const char *scale, unsigned idx) {
  Handle ds{H5Dopen2(file, dataset, H5P_DEFAULT), H5Dclose};
  Handle sc{H5Dopen2(file, scale, H5P_DEFAULT), H5Dclose};
  if (ds.id < 0 || sc.id < 0 || H5DSattach_scale(ds.id, sc.id, idx) < 0)
    return fail(MPAS_EINVAL, std::string("cannot attach ") + scale + " to " + dataset);
  return 0;
}

int mesh_field(hid_t file, const char *name, const void *data, int64_t rows, int64_t cols,
               int is_int, int space) {
  if (!data) return fail(MPAS_EINVAL, std::string("mpas_output_create: ") + name + " is NULL");
  if (int rc = write_dataset(file, mpas_io_dataset{name, data, rows, cols, is_int})) return rc;
  return attach(file, name, kSpace[space], 0);
}

}  // namespace

extern "C" {

int mpas_output_create(const char *path, int ntimes, const mpas_output_mesh *m,
                       mpas_output **out) {
  if (!path || !out || !m || ntimes < 1 || m->nCells < 1 || m->nEdges < 1 || m->nVertices < 1 ||
      m->vertexDegree < 1 || m->nVertLevels < 1 || m->nTracers < 0)
    return fail(MPAS_EINVAL, "mpas_output_create: bad argument");
  *out = nullptr;
  quiet();
  auto o = std::make_unique<mpas_output>();
  o->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (o->file < 0) return fail(MPAS_EINVAL, std::string("cannot create ") + path);
  o->ntimes = ntimes;
  o->n[0] = m->nCells;
  o->n[1] = m->nEdges;
  o->n[2] = m->nVertices;
  o->levels = m->nVertLevels;
  o->tracers = m->nTracers;
  {
    Handle sp{H5Screate(H5S_SCALAR), H5Sclose};
    Handle ty{H5Tcopy(H5T_C_S1), H5Tclose};
    H5Tset_size(ty.id, 3);
    H5Tset_strpad(ty.id, H5T_STR_NULLTERM);
    Handle at{H5Acreate2(o->file, "on_a_sphere", ty.id, sp.id, H5P_DEFAULT, H5P_DEFAULT),
              H5Aclose};
    H5Awrite(at.id, ty.id, "YES");
  }
  const hid_t f = o->file;
  const struct { const char *name; int64_t len; } scales[] = {
      {"nCells", m->nCells},           {"nEdges", m->nEdges},
      {"nVertices", m->nVertices},     {"vertexDegree", m->vertexDegree},
      {"nVertLevels", m->nVertLevels}, {"Time", ntimes},
      {"nTracers", m->nTracers}};
  for (const auto &s : scales)
    if (int rc = dim_scale(f, s.name, s.len)) return rc;
  std::vector<double> ones;
  const double *density = m->meshDensity;
  if (!density) {
    ones.assign(size_t(m->nCells), 1.0);
    density = ones.data();
  }
  const struct { const char *name; const double *data; int space; } fields[] = {
      {"areaCell", m->areaCell, 0},   {"xCell", m->xCell, 0},
      {"yCell", m->yCell, 0},         {"zCell", m->zCell, 0},
      {"latCell", m->latCell, 0},     {"lonCell", m->lonCell, 0},
      {"latVertex", m->latVertex, 2}, {"lonVertex", m->lonVertex, 2},
      {"mDensity", density, 0},       {"dvEdge", m->dvEdge, 1},
      {"dcEdge", m->dcEdge, 1}};
  for (const auto &x : fields)
    if (int rc = mesh_field(f, x.name, x.data, o->n[x.space], 1, 0, x.space)) return rc;
  if (int rc = mesh_field(f, "cellsOnVertex", m->cellsOnVertex, m->nVertices, m->vertexDegree,
                          1, 2))
    return rc;
  if (int rc = attach(f, "cellsOnVertex", "vertexDegree", 1)) return rc;
  *out = o.release();
  return 0;
}

int mpas_output_write(mpas_output *o, int t, const mpas_output_field *x) {
  if (!o || !x || !x->name || !x->data || x->space < 0 || x->space > 2 || x->levels < 0 ||
      x->tracers < 0 || (x->tracers && !x->levels))
    return fail(MPAS_EINVAL, "mpas_output_write: bad argument");
  if (t < 0 || t >= o->ntimes)
    return fail(MPAS_EINVAL, "mpas_output_write: time index " + std::to_string(t) +
                                 " outside [0, " + std::to_string(o->ntimes) + ")");
  if ((x->levels && x->levels != o->levels) || (x->tracers && x->tracers != o->tracers))
    return fail(MPAS_ESHAPE, std::string(x->name) + ": levels or tracers differ from the file's");
  int rank = 2;
  hsize_t dims[4] = {hsize_t(o->ntimes), hsize_t(o->n[x->space]), 1, 1};
  const char *scale[4] = {"Time", kSpace[x->space], "nVertLevels", "nTracers"};
  if (x->levels) dims[rank++] = hsize_t(x->levels);
  if (x->tracers) dims[rank++] = hsize_t(x->tracers);
  const bool fresh = !has_dataset(o->file, x->name);
  hid_t ds;
  if (fresh) {
    Handle sp{H5Screate_simple(rank, dims, nullptr), H5Sclose};
    ds = H5Dcreate2(o->file, x->name, H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT,
                    H5P_DEFAULT);
  } else {
    ds = H5Dopen2(o->file, x->name, H5P_DEFAULT);
  }
  {
    Handle d{ds, H5Dclose};
    if (d.id < 0) return fail(MPAS_EINVAL, std::string("cannot open dataset ") + x->name);
    Handle fs{H5Dget_space(d.id), H5Sclose};
    hsize_t have[4] = {0, 0, 0, 0};
    if (H5Sget_simple_extent_ndims(fs.id) != rank ||
        H5Sget_simple_extent_dims(fs.id, have, nullptr) != rank ||
        !std::equal(have + 1, have + rank, dims + 1))
      return fail(MPAS_ESHAPE, std::string(x->name) + ": shape differs from its first write");
    hsize_t start[4] = {hsize_t(t), 0, 0, 0}, count[4];
    std::copy(dims, dims + 4, count);
    count[0] = 1;
    H5Sselect_hyperslab(fs.id, H5S_SELECT_SET, start, nullptr, count, nullptr);
    Handle ms{H5Screate_simple(rank, count, nullptr), H5Sclose};
    if (H5Dwrite(d.id, H5T_NATIVE_DOUBLE, ms.id, fs.id, H5P_DEFAULT, x->data) < 0)
      return fail(MPAS_EINVAL, std::string("cannot write ") + x->name);
  }
  if (fresh) {  // scales attached at the first write (tasks/io.hh:24-29)
    if (int rc = attach(o->file, x->name, scale[0], 0)) return rc;
    if (int rc = attach(o->file, x->name, scale[1], 1)) return rc;
    for (int d = 2; d < rank; d++)
      if (int rc = attach(o->file, x->name, scale[d], unsigned(d))) return rc;
  }
  return 0;
}

int mpas_output_close(mpas_output *o) {
  if (!o) return fail(MPAS_EINVAL, "mpas_output_close: NULL");
  const hid_t f = o->file;
  o->file = -1;
  delete o;
  if (f >= 0 && H5Fclose(f) < 0) return fail(MPAS_EINVAL, "mpas_output_close: H5Fclose failed");
  return 0;
}

int mpas_output_read(const char *path, const char *name, int *rank, int64_t dims[4],
                     double *out, int64_t n) {
  if (!path || !name || !rank || !dims) return fail(MPAS_EINVAL, "mpas_output_read: NULL");
  quiet();
  Handle f{H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose};
  if (f.id < 0) return fail(MPAS_EINVAL, std::string("cannot open ") + path);
  if (!has_dataset(f.id, name)) return fail(MPAS_EINVAL, std::string("no dataset ") + name);
  Handle ds{H5Dopen2(f.id, name, H5P_DEFAULT), H5Dclose};
  Handle sp{H5Dget_space(ds.id), H5Sclose};
  const int r = H5Sget_simple_extent_ndims(sp.id);
  if (r < 0 || r > 4) return fail(MPAS_ESHAPE, std::string(name) + ": rank above 4");
  hsize_t d[4] = {1, 1, 1, 1};
  H5Sget_simple_extent_dims(sp.id, d, nullptr);
  *rank = r;
  int64_t total = 1;
  for (int i = 0; i < 4; i++) {
    dims[i] = i < r ? int64_t(d[i]) : 1;
    total *= dims[i];
  }
  if (!out) return 0;
  if (n != total)
    return fail(MPAS_ESHAPE, std::string(name) + " has " + std::to_string(total) +
                                 " entries, not " + std::to_string(n));
  if (H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0)
    return fail(MPAS_EINVAL, std::string("cannot read ") + name);
  return 0;
}

namespace {
herr_t scale_name_of(hid_t, unsigned, hid_t scale, void *data) {
  auto *s = static_cast<std::string *>(data);
  const ssize_t len = H5DSget_scale_name(scale, nullptr, 0);
  if (len > 0) {
    std::vector<char> b(size_t(len) + 1);
    H5DSget_scale_name(scale, b.data(), b.size());
    *s = b.data();
  }
  return 1;  // the first scale attached
}
}  // namespace

int mpas_output_scale_name(const char *path, const char *name, int dim, char *buf,
                           int64_t len) {
  if (!path || !name || !buf || len < 1) return fail(MPAS_EINVAL, "mpas_output_scale_name: NULL");
  quiet();
  Handle f{H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose};
  if (f.id < 0) return fail(MPAS_EINVAL, std::string("cannot open ") + path);
  if (!has_dataset(f.id, name)) return fail(MPAS_EINVAL, std::string("no dataset ") + name);
  Handle ds{H5Dopen2(f.id, name, H5P_DEFAULT), H5Dclose};
  std::string s;
  if (H5DSget_num_scales(ds.id, unsigned(dim)) > 0)
    H5DSiterate_scales(ds.id, unsigned(dim), nullptr, scale_name_of, &s);
  if (int64_t(s.size()) >= len) return fail(MPAS_EINVAL, "mpas_output_scale_name: buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return 0;
}

}  // extern "C"
