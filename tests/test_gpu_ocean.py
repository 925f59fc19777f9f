"""The tracer-only ocean driver (apps/ocean/run.cc:10-50) on the stage-fused
FCT kernels (mpas_ocean_rk4_step), against the oracle's restatement of the
same driver (per-task order, bit for bit)."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, bitwise_equal, mesh

pytestmark = pytest.mark.gpu


def _solid_body(m, u0):
    psi = -m.radius * u0 * np.sin(m["latVertex"])
    voe = m["verticesOnEdge"].reshape(-1, 2) - 1
    return -(psi[voe[:, 1]] - psi[voe[:, 0]]) / m["dvEdge"]


def _run_both(level, which, nT, steps, dt, kappa, nv=None, seed=0):
    from mpas_hip import TRACER_FCT, Context
    from oracle.oracle import OracleMesh, OracleOcean
    m = mesh(level)
    nv0, lt, t1 = m.ocean_case(which, a=m.radius, coef=kappa)
    nv = nv0 if nv is None else nv
    rng = np.random.default_rng(seed)
    cols = [t1] + [t1 * (1 + 0.1 * t) + rng.random(m.nCells) for t in range(1, nT)]
    tr = np.stack(cols, axis=1).ravel()
    om = OracleMesh(m)
    o = OracleOcean(om, nT=nT, mono=True, hmix=kappa != 0, hmix_coef=kappa, a=m.radius)
    o.init(lt, nv, tr)
    o.run(steps, dt)
    ref = o.tracers()
    c = Context(m, nTracers=nT, tracer_scheme=TRACER_FCT, kappa2=kappa)
    c.sw_upload(lt, nv, tr, np.zeros(m.nCells))
    c.ocean_rk4_step(dt, steps)
    h, u, got = c.sw_download()
    c.close()
    return ref, got, (lt, nv), (h, u)


@pytest.mark.parametrize("which,nT,kappa", [(1, 1, 0.0), (2, 2, 0.0), (1, 4, 1e3), (3, 3, 1e5)])
def test_ocean_driver_bitwise(which, nT, kappa):
    m = mesh(4)
    nv = _solid_body(m, 2 * np.pi * m.radius / 86400.0)
    ref, got, (lt, nv0), (h, u) = _run_both(4, which, nT, 6, 300.0, kappa, nv=nv)
    assert bitwise_equal(got, ref)
    assert bitwise_equal(h, lt) and bitwise_equal(u, nv)  # u and h stay fixed


def test_ocean_driver_case_flow_bitwise():
    """The test case's own velocity field (not divergence-free discretely)."""
    ref, got, _, _ = _run_both(3, 1, 2, 8, 600.0, 1e3)
    assert bitwise_equal(got, ref)


def test_ocean_driver_refuses_without_fct():
    from mpas_hip import TRACER_CENTRED, Context, MpasError
    m = mesh(3)
    c = Context(m, nTracers=1, tracer_scheme=TRACER_CENTRED)
    c.sw_upload(np.ones(m.nCells), np.zeros(m.nEdges), np.ones(m.nCells), np.zeros(m.nCells))
    with pytest.raises(MpasError):
        c.ocean_rk4_step(300.0, 1)
    c.close()


OCEAN = os.path.join(ROOT, "mpas-o-flecsi-2.0_amd", "bin", "ocean")


def _fnv1a(a):
    h = 1469598103934665603
    for b in np.ascontiguousarray(a, dtype=np.float64).tobytes():
        h ^= b
        h = (h * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return "%016x" % h


@pytest.mark.parametrize("tc", [1, 2, 3])
def test_ocean_executable_cases_against_oracle(tc):
    """bin/ocean runs apps/ocean's set-up, run and finalize.  Case 2 falls
    through into case 3 (advection and mixing on the harmonic, scored
    against the slotted cylinder); case 3 mixes only, which the executable
    runs as FCT at zero velocity: the oracle's advection-off run must match
    it bit for bit."""
    from oracle.oracle import OracleMesh, OracleOcean
    level, steps, dt, a = 4, 12, 96.0, 6371000.0
    r = subprocess.run([OCEAN, "--level", str(level), "-t", str(tc), "-s", str(steps), "-d",
                        str(dt)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    from mpas_hip import Mesh
    m = Mesh(level, radius=a, omega=7.29212e-5)
    coef = 0.0 if tc == 1 else 1e7
    nv, lt, t0 = m.ocean_case(1 if tc == 1 else 3, a=a, coef=coef)
    o = OracleOcean(OracleMesh(m), nT=1, mono=tc != 3, hmix=tc != 1, hmix_coef=coef, a=a)
    o.init(lt, nv, t0)
    o.run(steps, dt)
    ref = o.tracers()
    assert got["cells"] == m.nCells
    assert got["tracer_fnv1a"] == _fnv1a(ref)
    l2, mx = m.ocean_error(tc, ref, steps * dt, a=a, coef=coef)
    assert (got["l2_error"], got["max_error"]) == (l2, mx)
    if tc == 1:  # a twelfth of a revolution of the Gaussian hill: small error
        assert 0 < l2 < 1e-3
