"""The oracle's gather loops may run on several threads (ORC_PAR): each
value keeps its operations and order, so a threaded run equals the serial
one bit for bit; the reference's scatter loops stay serial."""
import numpy as np

from conftest import bitwise_equal, mesh

from oracle.oracle import OracleMesh, OracleOcean, OracleSW, set_threads


def _sw(m, om, threads, nT, scheme, steps, dt):
    from mpas_hip import sw_workload_state
    set_threads(threads)
    u, h, tr = sw_workload_state(m, nT, seed=8)
    o = OracleSW(om, nT=nT, tracer_adv=scheme, visc2=1e5, kappa2=1e3)
    o.set("h", h)
    o.set("u", u)
    if nT:
        o.set("tracers", tr)
    o.set("h_s", np.full(m.nCells, 4500.0))
    o.init_diagnostics()
    o.run(steps, dt)
    out = [o.get("h"), o.get("u")] + ([o.get("tracers")] if nT else [])
    set_threads(1)
    return out


def test_threaded_sw_equals_serial():
    m = mesh(5)
    om = OracleMesh(m)
    for nT, scheme in ((4, 2), (2, 1)):
        a = _sw(m, om, 1, nT, scheme, 3, 480.0)
        b = _sw(m, om, 8, nT, scheme, 3, 480.0)
        assert set_threads(0) == 1
        for x, y in zip(a, b):
            assert bitwise_equal(x, y), (nT, scheme)


def test_threaded_ocean_equals_serial():
    m = mesh(4)
    om = OracleMesh(m)
    outs = []
    for threads in (1, 8):
        set_threads(threads)
        nv, lt, t0 = m.ocean_case(1, a=m.radius, coef=1e3)
        o = OracleOcean(om, nT=1, mono=True, hmix=True, hmix_coef=1e3, a=6371000.0)
        o.init(lt, nv, t0)
        o.run(5, 300.0)
        outs.append(o.tracers())
    set_threads(1)
    assert bitwise_equal(outs[0], outs[1])
