"""ORACLE — test infrastructure only.

ctypes wrapper of _build/liboracle.so, the CPU restatement of the
reference's task bodies (see orc_internal.h).  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this, and
only as the checker or the timed CPU baseline; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MPAS_ORACLE_LIB points the checker at another build of the oracle (a
# scratch copy with a planted mistake, to see which pin catches it)
LIB = os.environ.get("MPAS_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")

_lib = None
P_f64 = ctypes.POINTER(ctypes.c_double)
P_i32 = ctypes.POINTER(ctypes.c_int32)
vp, d, i32, i64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_int64


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OrcParams(ctypes.Structure):
    _fields_ = [("nL", ctypes.c_int32), ("nT", ctypes.c_int32), ("tracer_adv", ctypes.c_int32),
                ("freeze_u", ctypes.c_int32), ("gravity", d), ("a", d), ("coef3rd", d),
                ("visc2", d), ("kappa2", d)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_mesh_new.restype = vp
        L.orc_mesh_new.argtypes = [i32] * 5
        L.orc_mesh_free.argtypes = [vp]
        L.orc_mesh_set_int.argtypes = [vp, ctypes.c_char_p, P_i32, i64]
        L.orc_mesh_set_dbl.argtypes = [vp, ctypes.c_char_p, P_f64, i64]
        L.orc_mesh_finalize.argtypes = [vp]
        L.orc_mesh_setup_fct.argtypes = [vp, i32, d]
        L.orc_migs.argtypes = [i32, P_f64, P_f64]
        L.orc_set_threads.argtypes = [i32]
        L.orc_set_threads(1)  # serial unless set_threads asks for more
        L.orc_mesh_cnx.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(P_i32),
                                   ctypes.POINTER(P_i32), ctypes.POINTER(i32)]
        L.orc_mesh_get.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(vp), ctypes.POINTER(i64),
                                   ctypes.POINTER(i32)]
        L.orc_sw_new.restype = vp
        L.orc_sw_new.argtypes = [vp, vp]
        L.orc_sw_free.argtypes = [vp]
        L.orc_sw_set.argtypes = [vp, ctypes.c_char_p, P_f64, i64]
        L.orc_sw_get.argtypes = [vp, ctypes.c_char_p, P_f64, i64]
        L.orc_sw_init_diagnostics.argtypes = [vp]
        L.orc_sw_run.argtypes = [vp, i32, d, P_f64]
        L.orc_sw_compute_error.restype = d
        L.orc_sw_compute_error.argtypes = [vp]
        L.orc_ocn_new.restype = vp
        L.orc_ocn_new.argtypes = [vp, vp, i32, i32, d]
        L.orc_ocn_free.argtypes = [vp]
        L.orc_ocn_init.argtypes = [vp, P_f64, P_f64, P_f64]
        L.orc_ocn_run.argtypes = [vp, i32, d, P_f64]
        L.orc_ocn_get_tracers.argtypes = [vp, P_f64]
        for name, n in [("orc_sw_compute_solve_diagnostics", 9), ("orc_sw_compute_tend", 8),
                        ("orc_sw_compute_scalar_tend", 4), ("orc_sw_del2_viscosity", 4),
                        ("orc_ocn_calc_normal_thickness_flux", 3),
                        ("orc_ocn_calc_provisional_layer_thicknesses", 5),
                        ("orc_ocn_low_order", 3), ("orc_ocn_high_order", 3),
                        ("orc_ocn_correct_and_rescale", 11), ("orc_ocn_accumulate_fct", 6),
                        ("orc_ocn_hmix_del2", 3)]:
            getattr(L, name).restype = None
        _lib = L
    return _lib


def set_threads(n):
    """Threads of the oracle's gather loops (the scatter loops stay serial,
    so every value keeps its bits); returns the count in use."""
    return lib().orc_set_threads(int(n))


def _f(a):
    return None if a is None else a.ctypes.data_as(P_f64)


def _check(rc, where):
    if rc != 0:
        raise RuntimeError(f"{where}: {lib().orc_last_error().decode()}")


class OracleMesh:
    """The oracle's flattened mesh (mesh::initialize restated) and its FCT setup."""

    INT = ["verticesOnCell", "edgesOnCell", "cellsOnCell", "verticesOnEdge", "edgesOnEdge"]
    DBL = ["areaCell", "dvEdge", "dcEdge", "areaTriangle", "kiteAreasOnVertex", "fVertex",
           "weightsOnEdge", "xCell", "yCell", "zCell", "xVertex", "yVertex", "zVertex", "latCell",
           "lonCell", "latEdge", "lonEdge", "latVertex", "lonVertex", "angleEdge", "meshDensity"]

    def __init__(self, mesh, nL=1, a=None, setup=True, maxLevelCell=None,
                 boundaryCell=None, fields=None):
        L = lib()
        if a is None:
            a = mesh.radius
        self.mesh, self.nL = mesh, nL
        self.nC, self.nE, self.nV = mesh.nCells, mesh.nEdges, mesh.nVertices
        self._h = L.orc_mesh_new(self.nC, self.nE, self.nV, mesh.maxEdges, mesh.maxEdges2)
        self._keep = []
        ints = dict(mesh.ints)
        dbls = dict(mesh.dbls)
        if fields:
            for k, v in fields.items():
                (ints if v.dtype == np.int32 else dbls)[k] = v
        for f in self.INT:
            a_ = np.ascontiguousarray(ints[f], dtype=np.int32)
            self._keep.append(a_)
            _check(L.orc_mesh_set_int(self._h, f.encode(), a_.ctypes.data_as(P_i32), a_.size), f)
        for f in self.DBL:
            a_ = np.ascontiguousarray(dbls[f], dtype=np.float64)
            self._keep.append(a_)
            _check(L.orc_mesh_set_dbl(self._h, f.encode(), _f(a_), a_.size), f)
        for name, arr in (("maxLevelCell", maxLevelCell), ("boundaryCell", boundaryCell)):
            if arr is not None:
                a_ = np.ascontiguousarray(arr, dtype=np.int32)
                self._keep.append(a_)
                _check(L.orc_mesh_set_int(self._h, name.encode(), a_.ctypes.data_as(P_i32),
                                          a_.size), name)
        _check(L.orc_mesh_finalize(self._h), "finalize")
        if setup:
            _check(L.orc_mesh_setup_fct(self._h, nL, a), "setup_fct")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_mesh_free(self._h)
            self._h = None

    def connectivity(self, which):
        off, idx, rows = P_i32(), P_i32(), i32()
        _check(lib().orc_mesh_cnx(self._h, which.encode(), ctypes.byref(off), ctypes.byref(idx),
                                  ctypes.byref(rows)), which)
        o = np.ctypeslib.as_array(off, (rows.value + 1,)).copy()
        n = int(o[-1])
        x = np.ctypeslib.as_array(idx, (max(n, 1),))[:n].copy()
        return o, x

    def setup_field(self, name):
        p, n, is_int = vp(), i64(), i32()
        _check(lib().orc_mesh_get(self._h, name.encode(), ctypes.byref(p), ctypes.byref(n),
                                  ctypes.byref(is_int)), name)
        ty = ctypes.c_int32 if is_int.value else ctypes.c_double
        arr = np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ty)), (n.value,))
        return arr.copy()

    # per-task calls on numpy arrays (reference layouts)
    def call(self, name, *args):
        conv = [vp(self._h)]
        for a_ in args:
            if isinstance(a_, np.ndarray):
                conv.append(_f(a_))
            elif isinstance(a_, (float, np.floating)):
                conv.append(d(float(a_)))
            elif isinstance(a_, (int, np.integer)):
                conv.append(ctypes.c_int(int(a_)))
            else:
                conv.append(a_)
        getattr(lib(), name)(*conv)


class OracleSW:
    """apps/sw/run.cc restated: the CPU reference of the fused RK4 run."""

    def __init__(self, omesh, nL=1, nT=0, tracer_adv=0, freeze_u=False, gravity=9.80616,
                 a=6371000.0, coef3rd=0.25, visc2=0.0, kappa2=0.0):
        self.om = omesh
        self.nL, self.nT = nL, nT
        p = OrcParams(nL, nT, tracer_adv, int(freeze_u), gravity, a, coef3rd, visc2, kappa2)
        self._h = lib().orc_sw_new(omesh._h, ctypes.byref(p))
        if not self._h:
            raise RuntimeError("orc_sw_new: " + lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_sw_free(self._h)
            self._h = None

    def _size(self, name):
        base = name.split(".")[0]
        C, E, V = self.om.nC, self.om.nE, self.om.nV
        L, T = self.nL, self.nT
        return {"h": C * L, "u": E * L, "h_edge": E * L, "h_vertex": V * L, "ke": C * L,
                "pv_edge": E * L, "tracers": C * L * T, "circulation": V * L,
                "vorticity": V * L, "pv_vertex": V * L, "tend_h": C * L, "tend_u": E * L,
                "tend_tracers": C * L * T, "h_s": C, "divergence": C * L}[base]

    def set(self, name, arr):
        a_ = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        _check(lib().orc_sw_set(self._h, name.encode(), _f(a_), a_.size), name)

    def get(self, name):
        out = np.empty(self._size(name))
        _check(lib().orc_sw_get(self._h, name.encode(), _f(out), out.size), name)
        return out

    def init_diagnostics(self):
        _check(lib().orc_sw_init_diagnostics(self._h), "init_diagnostics")

    def run(self, nsteps, dt):
        sec = d()
        _check(lib().orc_sw_run(self._h, nsteps, dt, ctypes.byref(sec)), "run")
        return sec.value

    def compute_error(self):
        return lib().orc_sw_compute_error(self._h)


class OracleOcean:
    """apps/ocean/run.cc restated (tracer-only FCT / del2 run)."""

    def __init__(self, omesh, nL=1, nT=1, mono=True, hmix=False, hmix_coef=0.0, a=6371000.0):
        self.om, self.nL, self.nT = omesh, nL, nT
        p = OrcParams(nL, nT, 2, 0, 9.80616, a, 0.25, 0.0, 0.0)
        self._h = lib().orc_ocn_new(omesh._h, ctypes.byref(p), int(mono), int(hmix), hmix_coef)
        if not self._h:
            raise RuntimeError("orc_ocn_new: " + lib().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_ocn_free(self._h)
            self._h = None

    def init(self, layerThickness, normalVelocity, tracers):
        lt = np.ascontiguousarray(layerThickness, dtype=np.float64)
        nv = np.ascontiguousarray(normalVelocity, dtype=np.float64)
        tr = np.ascontiguousarray(tracers, dtype=np.float64)
        _check(lib().orc_ocn_init(self._h, _f(lt), _f(nv), _f(tr)), "ocn_init")

    def run(self, nsteps, dt):
        sec = d()
        _check(lib().orc_ocn_run(self._h, nsteps, dt, ctypes.byref(sec)), "ocn_run")
        return sec.value

    def tracers(self):
        out = np.empty(self.om.nC * self.nL * self.nT)
        _check(lib().orc_ocn_get_tracers(self._h, _f(out)), "ocn_get_tracers")
        return out
