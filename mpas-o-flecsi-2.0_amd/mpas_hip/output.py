"""The shallow-water run's output file (apps/sw/run.cc:73-113) through
lib/libmpas_io.so.

- `output_steps(nsteps, num_output_times)`: the steps after which the run
  writes an output time (output_flagger, mpasoflecsi/io/state.hh:22-38).
- `OutputFile`: mesh_output_init (mpasoflecsi/tasks/io.cc:160-244), one
  field row per output time (tasks/io.hh:13-47), mesh_output_finalize.
- `run_sw(ctx, dt, nsteps, output)`: apps/sw/run.cc:77-106 on the fused
  step, copying the fields to the host only at output times.
- `read_output(path, name)`, `scale_name(path, name, dim)`: reading back."""
import ctypes
import os

import numpy as np

from . import _lib as L
from .meshfile import io

CELLS, EDGES, VERTICES = 0, 1, 2


class OutputMesh(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("nCells", "nEdges", "nVertices", "vertexDegree",
                                               "nVertLevels", "nTracers")] + \
               [(n, ctypes.c_void_p) for n in ("areaCell", "xCell", "yCell", "zCell", "latCell",
                                               "lonCell", "latVertex", "lonVertex", "meshDensity",
                                               "dvEdge", "dcEdge", "cellsOnVertex")]


class OutputField(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("data", ctypes.c_void_p), ("space", ctypes.c_int32),
                ("levels", ctypes.c_int32), ("tracers", ctypes.c_int32)]


_SIG = {
    "mpas_output_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int,
                                          ctypes.POINTER(OutputMesh),
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "mpas_output_write": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                         ctypes.POINTER(OutputField)]),
    "mpas_output_close": (ctypes.c_int, [ctypes.c_void_p]),
    "mpas_output_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p,
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p,
                                        ctypes.c_int64]),
    "mpas_output_scale_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                              ctypes.c_char_p, ctypes.c_int64]),
}


def _io():
    lib = io()
    if not getattr(lib, "_output_sigs", False):
        for name, (res, args) in _SIG.items():
            fn = getattr(lib, name)
            fn.restype, fn.argtypes = res, args
        lib._output_sigs = True
    return lib


def _check(rc, where):
    if rc != 0:
        raise L.MpasError(rc, where, io().mpas_io_last_error().decode())


def output_steps(nsteps, num_output_times):
    """Steps i in [1, nsteps] after which apps/sw/run.cc:100 writes an output
    time: output_flagger::output_step (mpasoflecsi/io/state.hh:22-38) over
    block_partition(nsteps, num_output_times - 1) (util/part.hh:28-53), i.e.
    i is written when it is the last index of the block that owns it.

    The partition covers 0..nsteps-1 while the loop counts 1..nsteps, so the
    last step is written only when nsteps // (num_output_times - 1) == 1.
    One output time writes only time 0 (the reference divides by zero)."""
    if num_output_times < 2 or nsteps < 1:
        return []
    n, P = nsteps, num_output_times - 1
    out = []
    for i in range(1, n + 1):
        owner = (P * (i + 1) - 1) // n  # block_partition::owner
        if (owner + 1) * n // P - 1 == i:  # block_partition::high(owner)
            out.append(i)
    return out


class OutputFile:
    """An output file of `num_output_times` times over `mesh` (a Mesh or a
    MeshFile: the fields of mesh_output_init are read from it)."""

    def __init__(self, path, mesh, num_output_times, nLevels=1, nTracers=0):
        self._o = None
        keep = {}

        def arr(name, dtype=np.float64):
            a = np.ascontiguousarray(mesh[name], dtype=dtype)
            keep[name] = a
            return a.ctypes.data

        m = OutputMesh(mesh.nCells, mesh.nEdges, mesh.nVertices, 3, nLevels, nTracers)
        for name in ("areaCell", "xCell", "yCell", "zCell", "latCell", "lonCell", "latVertex",
                     "lonVertex", "dvEdge", "dcEdge"):
            setattr(m, name, arr(name))
        has = getattr(mesh, "has", None)
        if has is None or has("meshDensity"):
            m.meshDensity = arr("meshDensity")
        m.cellsOnVertex = arr("cellsOnVertex", np.int32)
        o = ctypes.c_void_p()
        _check(_io().mpas_output_create(os.fsencode(path), int(num_output_times),
                                        ctypes.byref(m), ctypes.byref(o)), "mpas_output_create")
        self._o = o.value
        self.path = path
        self.num_output_times = int(num_output_times)
        self.nLevels, self.nTracers = nLevels, nTracers
        self.nVertices = mesh.nVertices

    def write_field(self, t, name, data, space, levels=0, tracers=0):
        """Row t of a field: data[entity](level)(tracer)."""
        a = np.ascontiguousarray(data, dtype=np.float64)
        f = OutputField(name.encode(), a.ctypes.data, space, levels, tracers)
        _check(_io().mpas_output_write(self._o, int(t), ctypes.byref(f)), "mpas_output_write")

    def write(self, t, h, h_s, vorticity, pv_vertex, tracers=None):
        """The fields of apps/sw/state.hh:66-70 at output time t."""
        L_ = self.nLevels
        self.write_field(t, "h", h, CELLS, L_)
        self.write_field(t, "h_s", h_s, CELLS)
        self.write_field(t, "vorticity", vorticity, VERTICES, L_)
        self.write_field(t, "potential_vorticity", pv_vertex, VERTICES, L_)
        if self.nTracers:
            self.write_field(t, "tracers", tracers, CELLS, L_, self.nTracers)

    def write_context(self, t, ctx):
        """Output time t from a context's current state.  At time 0 the
        vorticity and pv_vertex rows are zero: the reference writes them
        before its first diagnostics, as init_extra_fields left them
        (apps/sw/tasks/testcase.cc:16-35; run.cc:79-92)."""
        if t == 0:
            zero = np.zeros(self.nVertices * self.nLevels)
            vort = pv = zero
        else:
            vort, pv = ctx.sw_field("vorticity"), ctx.sw_field("pv_vertex")
        tr = ctx.sw_field("tracers") if self.nTracers else None
        self.write(t, ctx.sw_field("h"), ctx.sw_field("h_s"), vort, pv, tr)

    def close(self):
        if self._o:
            o, self._o = self._o, None
            _check(_io().mpas_output_close(o), "mpas_output_close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_sw(ctx, dt, nsteps, output=None):
    """apps/sw/run.cc:77-106 on the fused step: output time 0, then the
    steps, with the steps between output times in one mpas_rk4_step call
    and the fields copied to the host only at output times."""
    if output is None:
        ctx.rk4_step(dt, nsteps)
        return
    output.write_context(0, ctx)
    done = 0
    for t, i in enumerate(output_steps(nsteps, output.num_output_times), start=1):
        ctx.rk4_step(dt, i - done)
        done = i
        output.write_context(t, ctx)
    if nsteps > done:
        ctx.rk4_step(dt, nsteps - done)


def read_output(path, name):
    """A dataset of an output (or any HDF5) file as float64, in its shape."""
    rank = ctypes.c_int()
    dims = (ctypes.c_int64 * 4)()
    p, n = os.fsencode(path), name.encode()
    _check(_io().mpas_output_read(p, n, ctypes.byref(rank), dims, None, 0), "mpas_output_read")
    shape = tuple(int(dims[i]) for i in range(rank.value))
    out = np.empty(shape, dtype=np.float64)
    _check(_io().mpas_output_read(p, n, ctypes.byref(rank), dims, out.ctypes.data, out.size),
           "mpas_output_read")
    return out


def scale_name(path, name, dim):
    """The dimension scale attached to dimension `dim` of a dataset."""
    buf = ctypes.create_string_buffer(256)
    _check(_io().mpas_output_scale_name(os.fsencode(path), name.encode(), int(dim), buf,
                                        len(buf)), "mpas_output_scale_name")
    return buf.value.decode()
