"""Generated icosahedral SCVT meshes and seeded initial states
(host/meshgen.cc, host/initial.cc), in MPAS file conventions.

The reference's meshes (data/meshes/Sphere_240QU.nc, Sphere_480QU.nc) are
not available, so every workload runs on meshes generated here.
"""
import ctypes

import numpy as np

from ._lib import MESH_LIB, MeshDesc, P_f64, P_i32

ORDER_NATIVE, ORDER_HILBERT = 0, 1
EARTH_RADIUS = 6371220.0       # mesh radius of the workloads (m)
REF_A = 6371000.0              # mpas_constants::a (mpasoflecsi/common/constants.hh:28)
OMEGA_F = 7.292e-5             # f = 2 * 7.292e-5 * sin(lat) in the workloads
GRAVITY = 9.80616              # mpas_constants::gravity

INT_FIELDS = ["nEdgesOnCell", "edgesOnCell", "cellsOnCell", "verticesOnCell", "cellsOnEdge",
              "verticesOnEdge", "nEdgesOnEdge", "edgesOnEdge", "cellsOnVertex", "edgesOnVertex"]
DBL_FIELDS = ["xCell", "yCell", "zCell", "latCell", "lonCell", "xEdge", "yEdge", "zEdge",
              "latEdge", "lonEdge", "xVertex", "yVertex", "zVertex", "latVertex", "lonVertex",
              "areaCell", "areaTriangle", "kiteAreasOnVertex", "dcEdge", "dvEdge", "angleEdge",
              "weightsOnEdge", "meshDensity", "fVertex", "fEdge", "fCell"]

_mg = None


def _lib():
    global _mg
    if _mg is None:
        lib = ctypes.CDLL(MESH_LIB)
        lib.mg_generate_oriented.restype = ctypes.c_void_p
        lib.mg_generate_oriented.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_double, ctypes.c_int]
        lib.mg_free.argtypes = [ctypes.c_void_p]
        lib.mg_last_error.restype = ctypes.c_char_p
        lib.mg_dim.restype = ctypes.c_int64
        lib.mg_dim.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        lib.mg_int_field.restype = P_i32
        lib.mg_int_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                     ctypes.POINTER(ctypes.c_int64)]
        lib.mg_dbl_field.restype = P_f64
        lib.mg_dbl_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p,
                                     ctypes.POINTER(ctypes.c_int64)]
        d, u64 = ctypes.c_double, ctypes.c_uint64
        vp = ctypes.c_void_p
        lib.mg_init_sw_balanced.argtypes = [vp, d, d, d, d, d, d, d, u64, P_f64, P_f64, P_f64]
        lib.mg_init_tracer_blobs.argtypes = [vp, d, ctypes.c_int, d, d, u64, P_f64]
        lib.mg_init_sw_case1.argtypes = [vp, d, P_f64, P_f64]
        lib.mg_init_sw_case5.argtypes = [vp, d, d, d, P_f64, P_f64, P_f64, P_f64, P_f64,
                                         ctypes.c_int]
        lib.mg_init_sw_case6.argtypes = [vp, d, d, d, P_f64, P_f64]
        lib.mg_init_ocean_case.argtypes = [vp, ctypes.c_int, d, d, d, P_f64, P_f64, P_f64]
        lib.mg_ocean_error.argtypes = [vp, ctypes.c_int, d, d, d, P_f64, ctypes.c_int64,
                                       ctypes.POINTER(d), ctypes.POINTER(d)]
        _mg = lib
    return _mg


def _p(a, ty):
    return a.ctypes.data_as(ty) if a is not None else None


class Mesh:
    """A spherical Voronoi mesh: `ints` / `dbls` hold the MPAS fields."""

    def __init__(self, level, radius=EARTH_RADIUS, lloyd_iters=2, order=ORDER_HILBERT,
                 omega=OMEGA_F, rotate=True):
        lib = _lib()
        h = lib.mg_generate_oriented(level, radius, lloyd_iters, order, omega, int(bool(rotate)))
        if not h:
            raise RuntimeError("mesh generation failed: " + lib.mg_last_error().decode())
        self._h = h
        self.level, self.radius = level, radius
        self.nCells = lib.mg_dim(h, b"nCells")
        self.nEdges = lib.mg_dim(h, b"nEdges")
        self.nVertices = lib.mg_dim(h, b"nVertices")
        self.maxEdges = lib.mg_dim(h, b"maxEdges")
        self.maxEdges2 = lib.mg_dim(h, b"maxEdges2")
        self.ints, self.dbls = {}, {}
        n = ctypes.c_int64()
        for f in INT_FIELDS:
            p = lib.mg_int_field(h, f.encode(), ctypes.byref(n))
            self.ints[f] = np.ctypeslib.as_array(p, (n.value,)).copy()
        for f in DBL_FIELDS:
            p = lib.mg_dbl_field(h, f.encode(), ctypes.byref(n))
            self.dbls[f] = np.ctypeslib.as_array(p, (n.value,)).copy()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _mg is not None:
            _mg.mg_free(h)
            self._h = None

    def __getitem__(self, name):
        return self.ints[name] if name in self.ints else self.dbls[name]

    def desc(self, maxLevelCell=None, boundaryCell=None):
        """The mpas_mesh_desc of this mesh (keeps references to its arrays)."""
        d = MeshDesc()
        d.nCells, d.nEdges, d.nVertices = self.nCells, self.nEdges, self.nVertices
        d.maxEdges, d.maxEdges2, d.vertexDegree = self.maxEdges, self.maxEdges2, 3
        for f in ["verticesOnCell", "edgesOnCell", "cellsOnCell", "verticesOnEdge",
                  "edgesOnEdge"]:
            setattr(d, f, _p(self.ints[f], P_i32))
        for f in ["areaCell", "dvEdge", "dcEdge", "areaTriangle", "kiteAreasOnVertex", "fVertex",
                  "weightsOnEdge", "xCell", "yCell", "zCell", "xVertex", "yVertex", "zVertex",
                  "meshDensity"]:
            setattr(d, f, _p(self.dbls[f], P_f64))
        keep = []
        if maxLevelCell is not None:
            a = np.ascontiguousarray(maxLevelCell, dtype=np.int32)
            keep.append(a)
            d.maxLevelCell = _p(a, P_i32)
        if boundaryCell is not None:
            a = np.ascontiguousarray(boundaryCell, dtype=np.int32)
            keep.append(a)
            d.boundaryCell = _p(a, P_i32)
        d._keep = keep
        return d

    # ---- initial states ------------------------------------------------------
    def sw_balanced(self, u0=20.0, h_mean=5000.0, bump_amp=10.0, bump_radius=500e3, seed=1,
                    omega=OMEGA_F, gravity=GRAVITY):
        """Balanced zonal flow u0*cos(lat) + a seeded Gaussian h bump: (u, h)."""
        u = np.zeros(self.nEdges)
        h = np.zeros(self.nCells)
        rc = _lib().mg_init_sw_balanced(self._h, self.radius, omega, gravity, u0, h_mean,
                                        bump_amp, bump_radius, seed, _p(u, P_f64), _p(h, P_f64),
                                        None)
        if rc:
            raise RuntimeError("mg_init_sw_balanced failed")
        return u, h

    def tracer_blobs(self, seed, nblobs=8, rmin=300e3, rmax=1500e3):
        """Normalised sum of `nblobs` seeded Gaussian blobs, in [0, 1]."""
        t = np.zeros(self.nCells)
        if _lib().mg_init_tracer_blobs(self._h, self.radius, nblobs, rmin, rmax, seed,
                                       _p(t, P_f64)):
            raise RuntimeError("mg_init_tracer_blobs failed")
        return t

    def sw_case1(self, a=REF_A):
        u = np.zeros(self.nEdges)
        h = np.zeros(self.nCells)
        _lib().mg_init_sw_case1(self._h, a, _p(u, P_f64), _p(h, P_f64))
        return u, h

    def sw_case5(self, a=REF_A, omega=7.29212e-5, gravity=GRAVITY, ntr=2):
        u, h, hs = np.zeros(self.nEdges), np.zeros(self.nCells), np.zeros(self.nCells)
        f = np.zeros(self.nVertices)
        tr = np.zeros(self.nCells * ntr)
        _lib().mg_init_sw_case5(self._h, a, omega, gravity, _p(u, P_f64), _p(h, P_f64),
                                _p(hs, P_f64), _p(f, P_f64), _p(tr, P_f64), ntr)
        return u, h, hs, f, tr.reshape(self.nCells, ntr)

    def ocean_error(self, which, tracers, elapsed, a=REF_A, coef=0.0, stride=1):
        """(l2, max) of the ocean case's end-of-run report
        (calc_error_case_<which>, apps/ocean/tasks/testcase.cc:175-294) for
        tracers[c * stride], tracer 0 of level 0."""
        tr = np.ascontiguousarray(tracers, dtype=np.float64)
        if tr.size < (self.nCells - 1) * stride + 1:
            raise ValueError("tracers is too short for the stride")
        l2, mx = ctypes.c_double(), ctypes.c_double()
        if _lib().mg_ocean_error(self._h, which, a, coef, elapsed, _p(tr, P_f64), stride,
                                 ctypes.byref(l2), ctypes.byref(mx)):
            raise ValueError("mg_ocean_error: bad argument")
        return l2.value, mx.value

    def ocean_case(self, which, a=REF_A, coef=0.0, elapsed=0.0):
        nv, lt, tr = np.zeros(self.nEdges), np.zeros(self.nCells), np.zeros(self.nCells)
        _lib().mg_init_ocean_case(self._h, which, a, coef, elapsed, _p(nv, P_f64), _p(lt, P_f64),
                                  _p(tr, P_f64))
        return nv, lt, tr


def sw_workload_state(mesh, n_tracers, seed=1, u0=20.0, bump_amp=10.0):
    """Seeded state of the shallow-water workloads: balanced flow about a
    5,000 m mean with a 500 km Gaussian bump, and n_tracers blob tracers
    (seeds seed+1 ... seed+n_tracers).  Returns u [E], h [C], tracers [C, nT]."""
    u, h = mesh.sw_balanced(u0=u0, bump_amp=bump_amp, seed=seed)
    tr = np.zeros((mesh.nCells, n_tracers))
    for t in range(n_tracers):
        tr[:, t] = mesh.tracer_blobs(seed + 1 + t)
    return u, h, tr
