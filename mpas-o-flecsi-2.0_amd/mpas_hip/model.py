"""Host-side mirror of the reference's task interface over the C ABI.

`Context` plays the FleCSI mesh slot plus the field storage: it is created
from a mesh (the counterpart of mesh::initialize, mpasoflecsi/specialization/
mesh.cc:67-116), and each reference task is a method with the task's
argument names and order (apps/sw/tasks/solve.hh, apps/ocean/tasks/solve.hh)
taking `DeviceArray` fields.  Errors raise `MpasError` instead of aborting
(the reference's flog_fatal / flog_assert).
"""
import ctypes

import numpy as np

from . import _lib as L
from .mesh import GRAVITY, REF_A


class DeviceArray:
    """A float64 / int32 buffer in device memory owned by a Context."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        p = ctypes.c_void_p()
        L.check(L.hip().mpas_buf_alloc(ctx._h, max(1, self.nbytes), ctypes.byref(p)),
                "mpas_buf_alloc")
        self.ptr = p.value

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        if a.size != self.n:
            raise ValueError(f"size {a.size} != {self.n}")
        if self.n:
            L.check(L.hip().mpas_buf_upload(self.ctx._h, self.ptr, a.ctypes.data, self.nbytes),
                    "mpas_buf_upload")
        return self

    def numpy(self):
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            L.check(L.hip().mpas_buf_download(self.ctx._h, out.ctypes.data, self.ptr,
                                              self.nbytes), "mpas_buf_download")
        return out

    def free(self):
        if self.ptr and self.ctx._h:
            L.check(L.hip().mpas_buf_free(self.ctx._h, self.ptr), "mpas_buf_free")
        self.ptr = None


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, DeviceArray):
        return x.ptr
    return int(x)


SETUP_INT = {"maxLevelEdgeTop": 1, "nAdvCellsForEdge": 1, "advCellsForEdge": 16,
             "highOrderAdvMask": None}
SETUP_DBL = {"edgeSignOnCell": ("c", 8), "meshScalingDel2": ("e", 1), "derivTwo": ("e", 32),
             "advCoef": ("e", 16), "advCoef_3rd": ("e", 16)}


def _connectivity(hm, which):
    rows, count = ctypes.c_int64(), ctypes.c_int64()
    lib = L.hip()
    L.check(lib.mpas_mesh_connectivity(hm, which.encode(), None, None, ctypes.byref(rows),
                                       ctypes.byref(count)), "mpas_mesh_connectivity")
    off = np.zeros(rows.value + 1, dtype=np.int32)
    idx = np.zeros(max(1, count.value), dtype=np.int32)
    L.check(lib.mpas_mesh_connectivity(hm, which.encode(), off.ctypes.data_as(L.P_i32),
                                       idx.ctypes.data_as(L.P_i32), ctypes.byref(rows),
                                       ctypes.byref(count)), "mpas_mesh_connectivity")
    return off, idx[:count.value]


def _setup_field(hm, name, nC, nE, nL):
    if name in SETUP_INT:
        per = SETUP_INT[name] or nL
        out = np.zeros(nE * per, dtype=np.int32)
    else:
        where, per = SETUP_DBL[name]
        out = np.zeros((nC if where == "c" else nE) * per)
    L.check(L.hip().mpas_mesh_setup_field(hm, name.encode(), out.ctypes.data, out.size),
            "mpas_mesh_setup_field")
    return out


class HostMesh:
    """Host-only mesh preparation (mpas_host_mesh_create): flattening and
    FCT setup without a device."""

    def __init__(self, mesh, nLevels=1, nTracers=0, tracer_scheme=L.TRACER_NONE, kappa2=0.0,
                 sphere_radius=None, maxLevelCell=None, boundaryCell=None):
        if sphere_radius is None:
            sphere_radius = mesh.radius
        self._h = None
        self.nCells, self.nEdges, self.nLevels = mesh.nCells, mesh.nEdges, nLevels
        p = L.Params(nLevels, nTracers, tracer_scheme, 0, GRAVITY, sphere_radius, 0.25, 0.0,
                     kappa2)
        d = mesh.desc(maxLevelCell, boundaryCell)
        h = ctypes.c_void_p()
        L.check(L.hip().mpas_host_mesh_create(ctypes.byref(d), ctypes.byref(p), ctypes.byref(h)),
                "mpas_host_mesh_create")
        self._h = h.value

    def __del__(self):
        if getattr(self, "_h", None):
            L.hip().mpas_host_mesh_destroy(self._h)
            self._h = None

    def connectivity(self, which):
        return _connectivity(self._h, which)

    def setup_field(self, name):
        return _setup_field(self._h, name, self.nCells, self.nEdges, self.nLevels)

    def reordered(self, order):
        """mpas_host_mesh_reorder: this mesh renumbered (ORDER_*)."""
        h = ctypes.c_void_p()
        L.check(L.hip().mpas_host_mesh_reorder(self._h, int(order), ctypes.byref(h)),
                "mpas_host_mesh_reorder")
        o = HostMesh.__new__(HostMesh)
        o._h, o.nCells, o.nEdges, o.nLevels = h.value, self.nCells, self.nEdges, self.nLevels
        o.nVertices = getattr(self, "nVertices", None)
        return o

    def order(self, nVertices):
        """The original id of each cell, edge and vertex."""
        cs = np.zeros(self.nCells, dtype=np.int32)
        es = np.zeros(self.nEdges, dtype=np.int32)
        vs = np.zeros(nVertices, dtype=np.int32)
        L.check(L.hip().mpas_host_mesh_order(self._h, cs.ctypes.data_as(L.P_i32),
                                             es.ctypes.data_as(L.P_i32),
                                             vs.ctypes.data_as(L.P_i32)), "mpas_host_mesh_order")
        return cs, es, vs


class Context:
    def __init__(self, mesh, nLevels=1, nTracers=0, tracer_scheme=L.TRACER_NONE, freeze_u=False,
                 gravity=GRAVITY, sphere_radius=None, coef3rd=0.25, visc2=0.0, kappa2=0.0,
                 maxLevelCell=None, boundaryCell=None):
        """sphere_radius normalises positions in the FCT stencil setup; it
        defaults to the mesh radius."""
        self._h = None
        if sphere_radius is None:
            sphere_radius = mesh.radius
        self.mesh = mesh
        p = L.Params(nLevels, nTracers, tracer_scheme, int(bool(freeze_u)), gravity,
                     sphere_radius, coef3rd, visc2, kappa2)
        d = mesh.desc(maxLevelCell, boundaryCell)
        h = ctypes.c_void_p()
        L.check(L.hip().mpas_ctx_create(ctypes.byref(d), ctypes.byref(p), ctypes.byref(h)),
                "mpas_ctx_create")
        self._h = h.value
        self.params = p
        dims = np.zeros(5, dtype=np.int64)
        L.check(L.hip().mpas_ctx_dims(self._h, dims.ctypes.data_as(L.P_i64)), "mpas_ctx_dims")
        self.nCells, self.nEdges, self.nVertices, self.nLevels, self.nTracers = map(int, dims)

    def reordered(self, order):
        """mpas_ctx_create_reordered: a context on this mesh renumbered
        (ORDER_GIVEN, ORDER_RCM or ORDER_MORTON); its fields are in its own
        numbering (order() maps each entity to its original id)."""
        h = ctypes.c_void_p()
        L.check(L.hip().mpas_ctx_create_reordered(self._h, int(order), ctypes.byref(h)),
                "mpas_ctx_create_reordered")
        c = Context.__new__(Context)
        c._h, c.mesh, c.params = h.value, self.mesh, self.params
        c.nCells, c.nEdges, c.nVertices = self.nCells, self.nEdges, self.nVertices
        c.nLevels, c.nTracers = self.nLevels, self.nTracers
        return c

    def order(self):
        """The original id of each cell, edge and vertex."""
        cs = np.zeros(self.nCells, dtype=np.int32)
        es = np.zeros(self.nEdges, dtype=np.int32)
        vs = np.zeros(self.nVertices, dtype=np.int32)
        L.check(L.hip().mpas_ctx_order(self._h, cs.ctypes.data_as(L.P_i32),
                                       es.ctypes.data_as(L.P_i32), vs.ctypes.data_as(L.P_i32)),
                "mpas_ctx_order")
        return cs, es, vs

    def close(self):
        if self._h:
            L.hip().mpas_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- buffers -------------------------------------------------------------
    def alloc(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def to_device(self, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype).ravel()
        return DeviceArray(self, a.size, dtype).upload(a)

    def cells(self, per=1):
        return self.alloc(self.nCells * self.nLevels * per)

    def edges(self, per=1):
        return self.alloc(self.nEdges * self.nLevels * per)

    def vertices(self):
        return self.alloc(self.nVertices * self.nLevels)

    def sync(self):
        L.check(L.hip().mpas_ctx_sync(self._h), "mpas_ctx_sync")

    # ---- mesh queries ----------------------------------------------------------
    def connectivity(self, which):
        return _connectivity(L.hip().mpas_ctx_host_mesh(self._h), which)

    def setup_field(self, name):
        return _setup_field(L.hip().mpas_ctx_host_mesh(self._h), name, self.nCells, self.nEdges,
                            self.nLevels)

    # ---- tasks -----------------------------------------------------------------
    def _call(self, name, *args):
        fn = getattr(L.hip(), name)
        conv = [self._h] + [a if isinstance(a, float) else _ptr(a) for a in args]
        L.check(fn(*conv), name)

    # apps/sw/tasks/solve.hh
    def compute_solve_diagnostics(self, h, u, h_edge, h_vertex, circulation, vorticity, ke,
                                  pv_edge, pv_vertex):
        self._call("mpas_sw_compute_solve_diagnostics", h, u, h_edge, h_vertex, circulation,
                   vorticity, ke, pv_edge, pv_vertex)

    def init_provis(self, *fields):
        self._call("mpas_sw_init_provis", *fields)

    def setup_timestep(self, u_old, h_old, tracers_old, u_new, h_new, tracers_new):
        self._call("mpas_sw_setup_timestep", u_old, h_old, tracers_old, u_new, h_new, tracers_new)

    def compute_tend(self, h_edge, h, h_s, u, ke, pv_edge, tend_h, tend_u):
        self._call("mpas_sw_compute_tend", h_edge, h, h_s, u, ke, pv_edge, tend_h, tend_u)

    def compute_scalar_tend(self, h_edge, u, tracers, tracer_tend):
        self._call("mpas_sw_compute_scalar_tend", h_edge, u, tracers, tracer_tend)

    def compute_substep(self, u_provis, u_old, u_tend, h_provis, h_old, h_tend, tracers_provis,
                        tracers_old, tracers_tend, rkweight):
        self._call("mpas_sw_compute_substep", u_provis, u_old, u_tend, h_provis, h_old, h_tend,
                   tracers_provis, tracers_old, tracers_tend, float(rkweight))

    def accumulate_update(self, u_new, u_tend, h_new, h_tend, tracers_new, tracers_tend,
                          rkweight):
        self._call("mpas_sw_accumulate_update", u_new, u_tend, h_new, h_tend, tracers_new,
                   tracers_tend, float(rkweight))

    def finalize_timestep(self, u_old, u_new, h_new, tracers_new):
        self._call("mpas_sw_finalize_timestep", u_old, u_new, h_new, tracers_new)

    def timeshift(self, *fields):
        self._call("mpas_sw_timeshift", *fields)

    def compute_error(self, h, latCell, lonCell, elapsed, sphere_radius=REF_A):
        out = ctypes.c_double()
        L.check(L.hip().mpas_sw_compute_error(self._h, _ptr(h), _ptr(latCell), _ptr(lonCell),
                                              sphere_radius, elapsed, ctypes.byref(out)),
                "mpas_sw_compute_error")
        return out.value

    def del2_viscosity(self, u, vorticity, divergence, tend_u):
        self._call("mpas_sw_del2_viscosity", u, vorticity, divergence, tend_u)

    # apps/ocean/tasks/solve.hh
    def copy_tracers(self, tracers, tracersNew, tracersPrev):
        self._call("mpas_fct_copy_tracers", tracers, tracersNew, tracersPrev)

    def reinitialize_tracer_tendencies(self, tend):
        self._call("mpas_fct_reinitialize_tracer_tendencies", tend)

    def calc_normal_thickness_flux(self, layerThicknessEdge, normalVelocity, normalThicknessFlux):
        self._call("mpas_fct_calc_normal_thickness_flux", layerThicknessEdge, normalVelocity,
                   normalThicknessFlux)

    def calc_provisional_layer_thicknesses(self, layerThickness, normalThicknessFlux, hProv,
                                           hProvInv, hNewInv, dt):
        self._call("mpas_fct_calc_provisional_layer_thicknesses", layerThickness,
                   normalThicknessFlux, hProv, hProvInv, hNewInv, float(dt))

    def calc_horiz_tracer_advection_flux_mono_low_order(self, normalThicknessFlux, tracers,
                                                        lowOrderFlux):
        self._call("mpas_fct_low_order", normalThicknessFlux, tracers, lowOrderFlux)

    def calc_horiz_tracer_advection_flux_mono_high_order(self, normalThicknessFlux, tracers,
                                                         highOrderFlux):
        self._call("mpas_fct_high_order", normalThicknessFlux, tracers, highOrderFlux)

    def correct_and_rescale_high_order_mono_horiz_flux(self, hProvInv, layerThickness,
                                                       normalThicknessFlux, lowOrderFlux,
                                                       tracers, highOrderFlux, flxIn, flxOut,
                                                       tend, tracerMax, tracerMin, dt):
        self._call("mpas_fct_correct_and_rescale", hProvInv, layerThickness, normalThicknessFlux,
                   lowOrderFlux, tracers, highOrderFlux, flxIn, flxOut, tend, tracerMax,
                   tracerMin, float(dt))

    def accumulate_horiz_tracer_advection_tend_mono_FCT(self, hProvInv, layerThickness,
                                                        highOrderFlux, tracers, tracersProv,
                                                        tend, dt):
        self._call("mpas_fct_accumulate_tend", hProvInv, layerThickness, highOrderFlux, tracers,
                   tracersProv, tend, float(dt))

    def tracer_hmix_del2_tend(self, layerThicknessEdge, tracers, tend, coef):
        self._call("mpas_hmix_del2_tend", layerThicknessEdge, tracers, tend, float(coef))

    def step_tracers_forward(self, tend, tracers, tracersNew, tracersPrev, rk_weight,
                             rk_subweight):
        self._call("mpas_fct_step_tracers_forward", tend, tracers, tracersNew, tracersPrev,
                   float(rk_weight), float(rk_subweight))

    def update_tracers(self, tracersNew, tracers):
        self._call("mpas_fct_update_tracers", tracersNew, tracers)

    # ---- fused run (apps/sw/run.cc) --------------------------------------------
    def sw_upload(self, h, u, tracers=None, h_s=None):
        h = np.ascontiguousarray(h, dtype=np.float64).ravel()
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        if h_s is None:
            h_s = np.zeros(self.nCells)
        h_s = np.ascontiguousarray(h_s, dtype=np.float64).ravel()
        tr = None
        if self.nTracers:
            tr = np.ascontiguousarray(tracers, dtype=np.float64).ravel()
            if tr.size != self.nCells * self.nLevels * self.nTracers:
                raise ValueError("tracers has the wrong size")
        if h.size != self.nCells * self.nLevels or u.size != self.nEdges * self.nLevels:
            raise ValueError("h or u has the wrong size")
        L.check(L.hip().mpas_sw_state_upload(self._h, h.ctypes.data_as(L.P_f64),
                                             u.ctypes.data_as(L.P_f64),
                                             tr.ctypes.data_as(L.P_f64) if tr is not None else None,
                                             h_s.ctypes.data_as(L.P_f64)), "mpas_sw_state_upload")

    def state_check(self):
        """mpas_sw_state_check: the number of non-finite values in the
        resident h, u and tracers (0 when every value is finite)."""
        n = ctypes.c_int64(-1)
        rc = L.hip().mpas_sw_state_check(self._h, ctypes.byref(n))
        if rc not in (0, -6):
            L.check(rc, "mpas_sw_state_check")
        return n.value

    def sw_download(self):
        h = np.empty(self.nCells * self.nLevels)
        u = np.empty(self.nEdges * self.nLevels)
        tr = np.empty(self.nCells * self.nLevels * self.nTracers)
        L.check(L.hip().mpas_sw_state_download(self._h, h.ctypes.data_as(L.P_f64),
                                               u.ctypes.data_as(L.P_f64),
                                               tr.ctypes.data_as(L.P_f64) if tr.size else None),
                "mpas_sw_state_download")
        return h, u, tr

    FIELD_SIZES = {"h": "c", "u": "e", "tracers": "t", "h_edge": "e", "ke": "c", "pv_edge": "e",
                   "vorticity": "v", "pv_vertex": "v", "h_s": "c1", "divergence": "c"}

    def sw_field(self, name):
        kind = self.FIELD_SIZES.get(name, "c")  # unknown names are refused by the library
        n = {"c": self.nCells * self.nLevels, "e": self.nEdges * self.nLevels,
             "v": self.nVertices * self.nLevels, "c1": self.nCells,
             "t": self.nCells * self.nLevels * self.nTracers}[kind]
        out = np.empty(n)
        L.check(L.hip().mpas_sw_field_download(self._h, name.encode(),
                                               out.ctypes.data_as(L.P_f64), n),
                "mpas_sw_field_download")
        return out

    def profile(self, on=True):
        L.check(L.hip().mpas_prof_enable(self._h, int(bool(on))), "mpas_prof_enable")

    def profile_report(self):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        L.check(L.hip().mpas_prof_report(self._h, buf, len(buf)), "mpas_prof_report")
        return json.loads(buf.value.decode())

    def copy_probe(self, dst, src, nbytes):
        """Plain streaming device copy (bandwidth / counter calibration)."""
        L.check(L.hip().mpas_buf_copy(self._h, _ptr(dst), _ptr(src), int(nbytes)),
                "mpas_buf_copy")

    def rk4_step(self, dt, nsteps=1):
        L.check(L.hip().mpas_rk4_step(self._h, float(dt), int(nsteps)), "mpas_rk4_step")

    def rk4_step_noalloc(self, dt, nsteps, h, u, tracers, h_s):
        """mpas_rk4_step_noalloc: nsteps fused steps of the state in the
        device arrays h, u, tracers (advanced in place; h_s read), with the
        step's intermediates in the library's fixed scratch array."""
        L.check(L.hip().mpas_rk4_step_noalloc(self._h, float(dt), int(nsteps), _ptr(h), _ptr(u),
                                              _ptr(tracers), _ptr(h_s)),
                "mpas_rk4_step_noalloc")

    def ocean_rk4_step(self, dt, nsteps=1):
        """The tracer-only ocean driver (u, h fixed) on the uploaded state."""
        L.check(L.hip().mpas_ocean_rk4_step(self._h, float(dt), int(nsteps)),
                "mpas_ocean_rk4_step")

    # ---- tangent-linear model and adjoint (device arrays in, device arrays out) --
    def state_arrays(self):
        """Three new device arrays shaped like the state (u, h, tracers)."""
        return (self.alloc(self.nEdges * self.nLevels), self.alloc(self.nCells * self.nLevels),
                self.alloc(max(1, self.nCells * self.nLevels * self.nTracers)))

    def _trp(self, x):
        return _ptr(x) if self.nTracers else None

    def adj_init(self, checkpoint_every):
        L.check(L.hip().mpas_adj_init(self._h, int(checkpoint_every)), "mpas_adj_init")

    def adj_forward(self, dt, nsteps, x0, x1=(None, None, None)):
        """x0 = (u0, h0, tr0) device arrays; the final state goes to x1."""
        L.check(L.hip().mpas_adj_forward(self._h, float(dt), int(nsteps), _ptr(x0[0]),
                                         _ptr(x0[1]), self._trp(x0[2]), _ptr(x1[0]),
                                         _ptr(x1[1]), self._trp(x1[2])), "mpas_adj_forward")

    def adj_backward(self, xbar1, xbar0):
        L.check(L.hip().mpas_adj_backward(self._h, _ptr(xbar1[0]), _ptr(xbar1[1]),
                                          self._trp(xbar1[2]), _ptr(xbar0[0]), _ptr(xbar0[1]),
                                          self._trp(xbar0[2])), "mpas_adj_backward")

    def tlm_forward(self, dt, nsteps, x0, dx0, x1=(None, None, None), dx1=(None, None, None)):
        L.check(L.hip().mpas_tlm_forward(self._h, float(dt), int(nsteps), _ptr(x0[0]),
                                         _ptr(x0[1]), self._trp(x0[2]), _ptr(dx0[0]),
                                         _ptr(dx0[1]), self._trp(dx0[2]), _ptr(x1[0]),
                                         _ptr(x1[1]), self._trp(x1[2]), _ptr(dx1[0]),
                                         _ptr(dx1[1]), self._trp(dx1[2])), "mpas_tlm_forward")

    def adj_config(self, ad_block=256, tl_block=256, overlap=True, graph=False):
        L.check(L.hip().mpas_adj_config(self._h, int(ad_block), int(tl_block), int(bool(overlap)),
                                        int(bool(graph))), "mpas_adj_config")

    def adj_get_config(self):
        v = [ctypes.c_int() for _ in range(4)]
        L.check(L.hip().mpas_adj_get_config(self._h, *[ctypes.byref(x) for x in v]),
                "mpas_adj_get_config")
        return dict(zip(("ad_block", "tl_block", "overlap", "graph"), (x.value for x in v)))

    def adj_autotune(self, dt, nsteps):
        """mpas_adj_autotune from the resident state; returns the trials."""
        import json
        buf = ctypes.create_string_buffer(1 << 14)
        L.check(L.hip().mpas_adj_autotune(self._h, float(dt), int(nsteps), buf, len(buf)),
                "mpas_adj_autotune")
        return json.loads(buf.value.decode())

    def adj_frozen_forward(self, nsteps, x0, x1):
        L.check(L.hip().mpas_adj_frozen_forward(self._h, int(nsteps), _ptr(x0[0]), _ptr(x0[1]),
                                                self._trp(x0[2]), _ptr(x1[0]), _ptr(x1[1]),
                                                self._trp(x1[2])), "mpas_adj_frozen_forward")


class Batch:
    """An ensemble: member i runs on contexts[mesh_of[i]] with time step dt[i]
    (mpas_batch_*).  The contexts must outlive the batch."""

    def __init__(self, contexts, mesh_of, dt):
        self._b = None
        self.contexts = list(contexts)
        self.mesh_of = np.ascontiguousarray(mesh_of, dtype=np.int32)
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        if dt.size != self.mesh_of.size:
            raise ValueError("mesh_of and dt differ in length")
        hs = (ctypes.c_void_p * len(self.contexts))(*[c._h for c in self.contexts])
        b = ctypes.c_void_p()
        L.check(L.hip().mpas_batch_create(hs, len(self.contexts), int(self.mesh_of.size),
                                          self.mesh_of.ctypes.data_as(L.P_i32),
                                          dt.ctypes.data_as(L.P_f64), ctypes.byref(b)),
                "mpas_batch_create")
        self._b = b.value
        self.n = int(self.mesh_of.size)

    def close(self):
        if self._b:
            L.check(L.hip().mpas_batch_destroy(self._b), "mpas_batch_destroy")
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, member, h, u, tracers=None, h_s=None):
        c = self.contexts[self.mesh_of[member]]
        h = np.ascontiguousarray(h, dtype=np.float64).ravel()
        u = np.ascontiguousarray(u, dtype=np.float64).ravel()
        h_s = np.ascontiguousarray(np.zeros(c.nCells) if h_s is None else h_s,
                                   dtype=np.float64).ravel()
        tr = np.ascontiguousarray(tracers, dtype=np.float64).ravel() if c.nTracers else None
        L.check(L.hip().mpas_batch_upload(self._b, int(member), h.ctypes.data_as(L.P_f64),
                                          u.ctypes.data_as(L.P_f64),
                                          tr.ctypes.data_as(L.P_f64) if tr is not None else None,
                                          h_s.ctypes.data_as(L.P_f64)), "mpas_batch_upload")

    def download(self, member):
        c = self.contexts[self.mesh_of[member]]
        h = np.empty(c.nCells * c.nLevels)
        u = np.empty(c.nEdges * c.nLevels)
        tr = np.empty(c.nCells * c.nLevels * c.nTracers)
        L.check(L.hip().mpas_batch_download(self._b, int(member), h.ctypes.data_as(L.P_f64),
                                            u.ctypes.data_as(L.P_f64),
                                            tr.ctypes.data_as(L.P_f64) if tr.size else None),
                "mpas_batch_download")
        return h, u, tr

    def step(self, nsteps, mode=L.BATCH_GRAPH):
        L.check(L.hip().mpas_batch_step(self._b, int(nsteps), int(mode)), "mpas_batch_step")

    def sync(self):
        L.check(L.hip().mpas_batch_sync(self._b), "mpas_batch_sync")

    def groups(self):
        return L.check_count(L.hip().mpas_batch_groups(self._b), "mpas_batch_groups")

    def config(self, block=None, members_per_lane=1, order=0, xcd=1, wave_entity=1):
        blk = None
        if block is not None:
            blk = np.ascontiguousarray(block, dtype=np.int32)
        L.check(L.hip().mpas_batch_config(self._b, blk.ctypes.data_as(L.P_i32) if blk is not None
                                          else None, int(members_per_lane), int(order), int(xcd),
                                          int(wave_entity)), "mpas_batch_config")

    def get_config(self):
        blk = np.zeros(6, dtype=np.int32)
        lpl, o, x, we = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        L.check(L.hip().mpas_batch_get_config(self._b, blk.ctypes.data_as(L.P_i32),
                                              ctypes.byref(lpl), ctypes.byref(o),
                                              ctypes.byref(x), ctypes.byref(we)),
                "mpas_batch_get_config")
        return {"block": blk.tolist(), "members_per_lane": lpl.value, "order": o.value,
                "xcd": x.value, "wave_entity": we.value}

    def set_order(self, order):
        """mpas_batch_set_order: run the groups on reordered meshes (ORDER_*)."""
        L.check(L.hip().mpas_batch_set_order(self._b, int(order)), "mpas_batch_set_order")

    def get_order(self):
        return L.check_count(L.hip().mpas_batch_get_order(self._b), "mpas_batch_get_order")

    def tuned_mode(self):
        """The launch mode the autotuner found fastest (BATCH_GRAPH or
        BATCH_PERSISTENT)."""
        return L.check_count(L.hip().mpas_batch_tuned_mode(self._b), "mpas_batch_tuned_mode")

    def layout_blocks(self):
        """Workgroup sizes per kernel (rows) and group layout (columns), 0 for
        the kernel's own size."""
        b = np.zeros(36, dtype=np.int32)
        L.check(L.hip().mpas_batch_get_layout_blocks(self._b, b.ctypes.data_as(L.P_i32)),
                "mpas_batch_get_layout_blocks")
        return b.reshape(6, 6)

    def set_layout_blocks(self, blocks):
        b = np.ascontiguousarray(blocks, dtype=np.int32).ravel()
        if b.size != 36:
            raise ValueError("6 x 6 workgroup sizes")
        L.check(L.hip().mpas_batch_set_layout_blocks(self._b, b.ctypes.data_as(L.P_i32)),
                "mpas_batch_set_layout_blocks")

    def autotune(self, trial_steps=3):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        L.check(L.hip().mpas_batch_autotune(self._b, int(trial_steps), buf, len(buf)),
                "mpas_batch_autotune")
        return json.loads(buf.value.decode())
