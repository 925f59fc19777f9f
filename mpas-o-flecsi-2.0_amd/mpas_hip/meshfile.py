"""MPAS mesh files (HDF5 / netCDF-4 layout) through lib/libmpas_io.so.

`MeshFile(path)` presents the interface `Context` takes from `Mesh` (sizes,
`radius`, the MPAS fields by name, `desc()`); `write_mesh_file(path, mesh)`
writes a mesh's MPAS fields with their MPAS names and shapes."""
import ctypes
import os

import numpy as np

from . import _lib as L
from .mesh import DBL_FIELDS, INT_FIELDS

IO_LIB = os.path.join(L.LIB_DIR, "libmpas_io.so")


class Dataset(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("data", ctypes.c_void_p), ("rows", ctypes.c_int64),
                ("cols", ctypes.c_int64), ("is_int", ctypes.c_int32)]


_io = None


def io():
    """The loaded libmpas_io.so (raises if it has not been built)."""
    global _io
    if _io is None:
        if not os.path.exists(IO_LIB):
            raise ImportError(f"{IO_LIB} is missing: it is built when the HDF5 headers are "
                              "found (make -C mpas-o-flecsi-2.0_amd HDF5_DIR=...)")
        lib = ctypes.CDLL(IO_LIB)
        P = ctypes.c_void_p
        sig = {
            "mpas_io_last_error": (ctypes.c_char_p, []),
            "mpas_mesh_file_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(P)]),
            "mpas_mesh_file_close": (ctypes.c_int, [P]),
            "mpas_mesh_file_dims": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int64)]),
            "mpas_mesh_file_radius": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_double)]),
            "mpas_mesh_file_desc": (ctypes.c_int, [P, ctypes.POINTER(L.MeshDesc)]),
            "mpas_mesh_file_shape": (ctypes.c_int, [P, ctypes.c_char_p,
                                                    ctypes.POINTER(ctypes.c_int64)]),
            "mpas_mesh_file_read": (ctypes.c_int, [P, ctypes.c_char_p, P, ctypes.c_int64,
                                                   ctypes.c_int]),
            "mpas_mesh_file_write": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_double,
                                                    ctypes.c_int, ctypes.POINTER(Dataset)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        _io = lib
    return _io


def _check(rc, where):
    if rc != 0:
        raise L.MpasError(rc, where, io().mpas_io_last_error().decode())


# Columns of the 2-D MPAS fields (the rest are 1-D); a string names the
# mesh attribute that holds the width.
COLUMNS = {"edgesOnCell": "maxEdges", "verticesOnCell": "maxEdges", "cellsOnCell": "maxEdges",
           "cellsOnEdge": 2, "verticesOnEdge": 2, "edgesOnEdge": "maxEdges2",
           "weightsOnEdge": "maxEdges2", "cellsOnVertex": 3, "edgesOnVertex": 3,
           "kiteAreasOnVertex": 3}


class MeshFile:
    """A mesh read from an MPAS mesh file."""

    def __init__(self, path):
        self._f = None
        f = ctypes.c_void_p()
        _check(io().mpas_mesh_file_open(os.fsencode(path), ctypes.byref(f)), "mpas_mesh_file_open")
        self._f = f.value
        dims = (ctypes.c_int64 * 6)()
        _check(io().mpas_mesh_file_dims(self._f, dims), "mpas_mesh_file_dims")
        (self.nCells, self.nEdges, self.nVertices, self.maxEdges, self.maxEdges2,
         self.vertexDegree) = (int(x) for x in dims)
        r = ctypes.c_double()
        _check(io().mpas_mesh_file_radius(self._f, ctypes.byref(r)), "mpas_mesh_file_radius")
        self.radius = r.value
        self.path = path

    def close(self):
        if self._f:
            io().mpas_mesh_file_close(self._f)
            self._f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def has(self, name):
        sh = (ctypes.c_int64 * 2)()
        return io().mpas_mesh_file_shape(self._f, name.encode(), sh) == 0

    def __getitem__(self, name):
        sh = (ctypes.c_int64 * 2)()
        _check(io().mpas_mesh_file_shape(self._f, name.encode(), sh), "mpas_mesh_file_shape")
        n = int(sh[0] * sh[1])
        is_int = name in INT_FIELDS
        out = np.empty(n, dtype=np.int32 if is_int else np.float64)
        _check(io().mpas_mesh_file_read(self._f, name.encode(), out.ctypes.data, n, int(is_int)),
               "mpas_mesh_file_read")
        return out

    def desc(self, maxLevelCell=None, boundaryCell=None):
        """The mpas_mesh_desc of the file's arrays (valid while this object lives)."""
        d = L.MeshDesc()
        _check(io().mpas_mesh_file_desc(self._f, ctypes.byref(d)), "mpas_mesh_file_desc")
        keep = [self]
        if maxLevelCell is not None:
            a = np.ascontiguousarray(maxLevelCell, dtype=np.int32)
            keep.append(a)
            d.maxLevelCell = a.ctypes.data_as(L.P_i32)
        if boundaryCell is not None:
            a = np.ascontiguousarray(boundaryCell, dtype=np.int32)
            keep.append(a)
            d.boundaryCell = a.ctypes.data_as(L.P_i32)
        d._keep = keep
        return d


def write_mesh_file(path, mesh, fields=None):
    """Writes the MPAS fields of `mesh` (a Mesh or any object with the same
    sizes and `mesh[name]`), 1-based connectivity as MPAS files hold it."""
    rows_of = {"Cell": mesh.nCells, "Edge": mesh.nEdges, "Vertex": mesh.nVertices}
    special = {"areaTriangle": mesh.nVertices, "meshDensity": mesh.nCells}
    names = fields if fields is not None else INT_FIELDS + DBL_FIELDS
    keep, ds = [], []
    for name in names:
        cols = COLUMNS.get(name, 1)
        if isinstance(cols, str):
            cols = getattr(mesh, cols)
        rows = special.get(name)
        if rows is None:
            rows = next(v for k, v in rows_of.items() if name.endswith(k))
        is_int = name in INT_FIELDS
        a = np.ascontiguousarray(mesh[name], dtype=np.int32 if is_int else np.float64)
        if a.size != rows * cols:
            raise ValueError(f"{name}: {a.size} entries for {rows} x {cols}")
        keep.append(a)
        ds.append(Dataset(name.encode(), a.ctypes.data, rows, cols, int(is_int)))
    arr = (Dataset * len(ds))(*ds)
    _check(io().mpas_mesh_file_write(os.fsencode(path), float(mesh.radius), len(ds), arr),
           "mpas_mesh_file_write")
