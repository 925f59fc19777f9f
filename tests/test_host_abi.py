"""CPU checks of the C ABI library: it loads, exports what include/mpas_hip.h
declares, prepares the mesh on the host exactly as the oracle does, and
rejects bad input with status codes (no device needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, mesh

HEADER = os.path.join(ROOT, "include", "mpas_hip.h")


def _declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mpas_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from mpas_hip import hip
    lib = hip()
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_the_header():
    from mpas_hip._lib import SIGNATURES
    assert set(_declared()) <= set(SIGNATURES), set(_declared()) - set(SIGNATURES)


@pytest.mark.parametrize("level", [3, 4])
def test_host_mesh_bitwise_equal_to_oracle(level):
    from mpas_hip import TRACER_FCT, HostMesh
    from oracle.oracle import OracleMesh
    m = mesh(level)
    hm = HostMesh(m, nTracers=2, tracer_scheme=TRACER_FCT)
    om = OracleMesh(m)
    for w in ["c2v", "c2e", "c2c", "e2v", "e2e", "v2c", "e2c", "v2e"]:
        a, b = hm.connectivity(w), om.connectivity(w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), w
    for f in ["maxLevelEdgeTop", "nAdvCellsForEdge", "advCellsForEdge", "highOrderAdvMask",
              "edgeSignOnCell", "meshScalingDel2", "derivTwo", "advCoef", "advCoef_3rd"]:
        a, b = hm.setup_field(f), om.setup_field(f)
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), f


def test_host_mesh_partial_levels_match_oracle(mesh3):
    from mpas_hip import TRACER_FCT, HostMesh
    from oracle.oracle import OracleMesh
    m, nL = mesh3, 3
    rng = np.random.default_rng(5)
    mlc = rng.integers(1, nL + 1, size=m.nCells).astype(np.int32)
    bnd = (rng.random(m.nCells * nL) < 0.2).astype(np.int32)
    hm = HostMesh(m, nLevels=nL, nTracers=1, tracer_scheme=TRACER_FCT, maxLevelCell=mlc,
                  boundaryCell=bnd)
    om = OracleMesh(m, nL=nL, maxLevelCell=mlc, boundaryCell=bnd)
    for f in ["maxLevelEdgeTop", "highOrderAdvMask"]:
        assert np.array_equal(hm.setup_field(f), om.setup_field(f)), f


def test_level8_setup_bitwise(request):
    """The benchmark mesh: the stencil setup of the product equals the oracle's."""
    from mpas_hip import TRACER_FCT, HostMesh
    from oracle.oracle import OracleMesh
    m = mesh(8)
    hm = HostMesh(m, nTracers=4, tracer_scheme=TRACER_FCT, kappa2=1e3)
    om = OracleMesh(m)
    for f in ["advCoef", "advCoef_3rd", "nAdvCellsForEdge"]:
        a, b = hm.setup_field(f), om.setup_field(f)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f
        assert np.isfinite(a.astype(np.float64)).all()


# ---- error behaviour ----------------------------------------------------------

def _desc_with(m, **override):
    """A mesh_desc whose arrays may be replaced (kept alive on the object)."""
    d = m.desc()
    keep = []
    from mpas_hip._lib import P_f64, P_i32
    for k, v in override.items():
        if v is None:
            setattr(d, k, None)
            continue
        keep.append(v)
        setattr(d, k, v.ctypes.data_as(P_i32 if v.dtype == np.int32 else P_f64))
    d._keep2 = keep
    return d


def _create(d, nT=0, scheme=0, kappa2=0.0, radius=None, nL=1):
    from mpas_hip import hip
    from mpas_hip._lib import Params
    p = Params(nL, nT, scheme, 0, 9.80616, radius or 6371220.0, 0.25, 0.0, kappa2)
    h = ctypes.c_void_p()
    rc = hip().mpas_host_mesh_create(ctypes.byref(d), ctypes.byref(p), ctypes.byref(h))
    if rc == 0:
        hip().mpas_host_mesh_destroy(h)
    return rc, hip().mpas_last_error().decode()


def test_rejects_out_of_range_ids(mesh3):
    m = mesh3
    eoc = m["edgesOnCell"].copy()
    eoc[5] = m.nEdges + 7
    rc, msg = _create(_desc_with(m, edgesOnCell=eoc))
    assert rc == -1 and "out of range" in msg


def test_rejects_open_mesh(mesh3):
    m = mesh3
    eoc = m["edgesOnCell"].reshape(-1, 8).copy()
    eoc[0, 0] = 0  # cell 0 loses an edge: that edge now has one cell
    rc, msg = _create(_desc_with(m, edgesOnCell=eoc.ravel()))
    assert rc == -1 and "cells" in msg


def test_rejects_null_and_bad_params(mesh3):
    m = mesh3
    rc, msg = _create(_desc_with(m, weightsOnEdge=None))
    assert rc == -1 and "NULL" in msg
    rc, msg = _create(m.desc(), nT=9)
    assert rc == -1
    rc, msg = _create(m.desc(), nT=0, scheme=2)
    assert rc == -1 and "nTracers" in msg
    rc, msg = _create(m.desc(), nT=1, scheme=2, radius=-1.0)
    assert rc == -1 and "radius" in msg


def test_rejects_mismatched_stencil_radius():
    """With a setup radius far from the mesh's the reference's fit divides
    0 by 0; the library refuses instead of producing NaN coefficients."""
    m = mesh(7)
    rc, msg = _create(m.desc(), nT=1, scheme=2, radius=6371000.0)
    assert rc == -1 and "not finite" in msg


def test_context_without_device_reports_hip_error(mesh3):
    from mpas_hip import Context, MpasError, hip
    if hip().mpas_ctx_create and os.environ.get("HIP_VISIBLE_DEVICES", "") == "":
        try:
            ctx = Context(mesh3)
        except MpasError as e:
            assert e.code == -2
        else:  # a device is present
            ctx.close()
