"""Parity of the HIP kernels with the oracle (the CPU restatement of the
reference's task bodies), through the C ABI.

Tolerance: the workloads require 1e-12 relative per step (FP64); the kernels
form every value with the oracle's operations in the oracle's order, so the
tests also assert bit equality where that is the design claim.
"""
import numpy as np
import pytest

from conftest import bitwise_equal, max_rel, mesh

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative, per step


def _rng_fields(m, nL, nT, seed):
    rng = np.random.default_rng(seed)
    C, E, V = m.nCells, m.nEdges, m.nVertices
    return dict(
        h=5000.0 + 200.0 * rng.standard_normal(C * nL),
        u=20.0 * rng.standard_normal(E * nL),
        h_s=4500.0 + 50.0 * rng.standard_normal(C),
        tr=rng.random(C * nL * nT),
        h_edge=5000.0 + 200.0 * rng.standard_normal(E * nL),
        ke=100.0 * rng.random(C * nL),
        pv_edge=1e-8 * rng.standard_normal(E * nL),
        vort=1e-5 * rng.standard_normal(V * nL),
    )


def _ctx(m, **kw):
    from mpas_hip import Context
    return Context(m, **kw)


def _omesh(m, nL=1, **kw):
    from oracle.oracle import OracleMesh
    return OracleMesh(m, nL=nL, **kw)


# ---- mesh preparation on the device context --------------------------------

def test_ctx_connectivity_matches_oracle(mesh4):
    from mpas_hip import TRACER_FCT
    ctx = _ctx(mesh4, nTracers=2, tracer_scheme=TRACER_FCT)
    om = _omesh(mesh4)
    for w in ["c2v", "c2e", "c2c", "e2v", "e2e", "v2c", "e2c", "v2e"]:
        a, b = ctx.connectivity(w), om.connectivity(w)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), w
    ctx.close()


# ---- compute_solve_diagnostics / compute_tend / compute_scalar_tend ---------

@pytest.mark.parametrize("nL", [1, 3])
def test_sw_tasks_bitwise(mesh3, nL):
    from mpas_hip import TRACER_CENTRED
    m = mesh3
    nT = 3
    C, E, V = m.nCells, m.nEdges, m.nVertices
    f = _rng_fields(m, nL, nT, seed=11 + nL)
    ctx = _ctx(m, nLevels=nL, nTracers=nT, tracer_scheme=TRACER_CENTRED, visc2=1e5)
    om = _omesh(m, nL=nL)
    d = {k: ctx.to_device(v) for k, v in f.items()}
    outs = {n: ctx.alloc(s) for n, s in [("h_edge", E * nL), ("h_vertex", V * nL),
                                         ("circ", V * nL), ("vort", V * nL), ("ke", C * nL),
                                         ("pv_edge", E * nL), ("pv_vertex", V * nL)]}
    ctx.compute_solve_diagnostics(d["h"], d["u"], outs["h_edge"], outs["h_vertex"], outs["circ"],
                                  outs["vort"], outs["ke"], outs["pv_edge"], outs["pv_vertex"])
    ref = {n: np.zeros(o.n) for n, o in outs.items()}
    om.call("orc_sw_compute_solve_diagnostics", nL, f["h"], f["u"], ref["h_edge"],
            ref["h_vertex"], ref["circ"], ref["vort"], ref["ke"], ref["pv_edge"],
            ref["pv_vertex"])
    for n in outs:
        assert bitwise_equal(outs[n].numpy(), ref[n]), (n, max_rel(outs[n].numpy(), ref[n]))

    tend_h, tend_u = ctx.alloc(C * nL), ctx.alloc(E * nL)
    ctx.compute_tend(d["h_edge"], d["h"], d["h_s"], d["u"], d["ke"], d["pv_edge"], tend_h, tend_u)
    r_th, r_tu = np.zeros(C * nL), np.zeros(E * nL)
    om.call("orc_sw_compute_tend", nL, 9.80616, f["h_edge"], f["h"], f["h_s"], f["u"], f["ke"],
            f["pv_edge"], r_th, r_tu)
    assert bitwise_equal(tend_h.numpy(), r_th), max_rel(tend_h.numpy(), r_th)
    assert bitwise_equal(tend_u.numpy(), r_tu), max_rel(tend_u.numpy(), r_tu)

    # del2 viscosity on u (adds to tend_u)
    div = ctx.alloc(C * nL)
    ctx.del2_viscosity(d["u"], d["vort"], div, tend_u)
    r_div = np.zeros(C * nL)
    om.call("orc_sw_del2_viscosity", nL, 1e5, f["u"], f["vort"], r_div, r_tu)
    assert bitwise_equal(div.numpy(), r_div)
    assert bitwise_equal(tend_u.numpy(), r_tu), max_rel(tend_u.numpy(), r_tu)

    tt = ctx.alloc(C * nL * nT)
    ctx.compute_scalar_tend(d["h_edge"], d["u"], d["tr"], tt)
    r_tt = np.zeros(C * nL * nT)
    om.call("orc_sw_compute_scalar_tend", nL, nT, f["h_edge"], f["u"], f["tr"], r_tt)
    assert bitwise_equal(tt.numpy(), r_tt), max_rel(tt.numpy(), r_tt)
    ctx.close()


# ---- FCT and del2 tracer tasks (apps/ocean/tasks/solve.cc) ------------------

@pytest.mark.parametrize("nL,partial", [(1, False), (3, False), (3, True)])
def test_fct_tasks_bitwise(mesh3, nL, partial):
    from mpas_hip import TRACER_FCT
    m = mesh3
    nT = 4
    C, E = m.nCells, m.nEdges
    rng = np.random.default_rng(7 + nL + 10 * partial)
    mlc = rng.integers(1, nL + 1, size=C).astype(np.int32) if partial else None
    bnd = (rng.random(C * nL) < 0.1).astype(np.int32) if partial else None
    f = _rng_fields(m, nL, nT, seed=3 + nL)
    ctx = _ctx(m, nLevels=nL, nTracers=nT, tracer_scheme=TRACER_FCT, kappa2=1e3,
               maxLevelCell=mlc, boundaryCell=bnd)
    om = _omesh(m, nL=nL, maxLevelCell=mlc, boundaryCell=bnd)
    for name in ["maxLevelEdgeTop", "highOrderAdvMask", "advCoef", "advCoef_3rd"]:
        assert bitwise_equal(ctx.setup_field(name).astype(np.float64),
                             om.setup_field(name).astype(np.float64)), name
    dt = 300.0
    lt, lte, nv, tr = f["h"], f["h_edge"], f["u"], f["tr"]
    d_lt, d_lte, d_nv, d_tr = (ctx.to_device(x) for x in (lt, lte, nv, tr))

    ntf = ctx.alloc(E * nL)
    ctx.calc_normal_thickness_flux(d_lte, d_nv, ntf)
    r_ntf = np.zeros(E * nL)
    om.call("orc_ocn_calc_normal_thickness_flux", nL, lte, nv, r_ntf)
    assert bitwise_equal(ntf.numpy(), r_ntf)

    init = 0.125  # values of the levels the tasks do not write
    hp, hpi, hni = (ctx.to_device(np.full(C * nL, init)) for _ in range(3))
    r_hp, r_hpi, r_hni = (np.full(C * nL, init) for _ in range(3))
    ctx.calc_provisional_layer_thicknesses(d_lt, ntf, hp, hpi, hni, dt)
    om.call("orc_ocn_calc_provisional_layer_thicknesses", nL, lt, r_ntf, r_hp, r_hpi, r_hni, dt)
    for a, b in ((hp, r_hp), (hpi, r_hpi), (hni, r_hni)):
        assert bitwise_equal(a.numpy(), b)

    lo, ho = ctx.alloc(E * nL * nT), ctx.alloc(E * nL * nT)
    ctx.calc_horiz_tracer_advection_flux_mono_low_order(ntf, d_tr, lo)
    ctx.calc_horiz_tracer_advection_flux_mono_high_order(ntf, d_tr, ho)
    r_lo, r_ho = np.zeros(E * nL * nT), np.zeros(E * nL * nT)
    om.call("orc_ocn_low_order", nL, nT, r_ntf, tr, r_lo)
    om.call("orc_ocn_high_order", nL, nT, 0.25, r_ntf, tr, r_ho)
    assert bitwise_equal(lo.numpy(), r_lo)
    assert bitwise_equal(ho.numpy(), r_ho), max_rel(ho.numpy(), r_ho)

    n = C * nL * nT
    tend = ctx.to_device(np.zeros(n))
    fin, fout, tmax, tmin = (ctx.to_device(np.full(n, init)) for _ in range(4))
    r_tend = np.zeros(n)
    r_fin, r_fout, r_tmax, r_tmin = (np.full(n, init) for _ in range(4))
    ctx.correct_and_rescale_high_order_mono_horiz_flux(hpi, d_lt, ntf, lo, d_tr, ho, fin, fout,
                                                       tend, tmax, tmin, dt)
    om.call("orc_ocn_correct_and_rescale", nL, nT, r_hpi, lt, r_ntf, r_lo, tr, r_ho, r_fin,
            r_fout, r_tend, r_tmax, r_tmin, dt)
    for a, b, nm in ((ho, r_ho, "ho"), (fin, r_fin, "flxIn"), (fout, r_fout, "flxOut"),
                     (tend, r_tend, "tend"), (tmax, r_tmax, "max"), (tmin, r_tmin, "min")):
        assert bitwise_equal(a.numpy(), b), (nm, max_rel(a.numpy(), b))

    trp = ctx.to_device(np.full(n, init))
    r_trp = np.full(n, init)
    ctx.accumulate_horiz_tracer_advection_tend_mono_FCT(hpi, d_lt, ho, d_tr, trp, tend, dt)
    om.call("orc_ocn_accumulate_fct", nL, nT, r_hpi, lt, r_ho, tr, r_trp, r_tend, dt)
    assert bitwise_equal(tend.numpy(), r_tend)
    assert bitwise_equal(trp.numpy(), r_trp)

    ctx.tracer_hmix_del2_tend(d_lte, d_tr, tend, 1e3)
    om.call("orc_ocn_hmix_del2", nL, nT, lte, tr, r_tend, 1e3)
    assert bitwise_equal(tend.numpy(), r_tend), max_rel(tend.numpy(), r_tend)
    ctx.close()


# ---- the fused RK4 step against the oracle's driver -------------------------

FUSED_CASES = {
    # name: (level, nL, nT, scheme, visc2, kappa2, freeze_u, dt, steps)
    "sw_visc": (4, 1, 0, 0, 1e5, 0.0, False, 960.0, 5),
    "sw_fct4": (3, 1, 4, 2, 1e5, 1e3, False, 600.0, 5),
    # the most tracers the fused step takes (two lanes of four per edge in
    # the high-order flux)
    "sw_fct8": (3, 1, 8, 2, 1e5, 1e3, False, 600.0, 3),
    # odd tracer counts: the momentum and flux phases in one launch, the
    # tracer update (gathers) in a launch of its own
    "sw_fct1": (3, 1, 1, 2, 1e5, 1e3, False, 600.0, 4),
    "sw_fct3": (4, 1, 3, 2, 1e5, 1e3, False, 480.0, 3),
    "sw_centred2": (3, 1, 2, 1, 0.0, 1e3, False, 600.0, 4),
    # four centred tracers: the cells kernel's two tracer-pair lanes share
    # the cell's edge rows
    "sw_centred4": (3, 1, 4, 1, 1e5, 1e3, False, 600.0, 3),
    "sw_fct_2levels": (3, 2, 2, 2, 1e5, 0.0, False, 600.0, 3),
    "case1_frozen": (3, 1, 1, 1, 0.0, 0.0, True, 40.0, 5),
    # one level per lane with many levels (the layout of multi-level contexts)
    "sw_fct_8levels": (3, 8, 2, 2, 1e5, 1e3, False, 600.0, 3),
    # 32 levels: level pairs put an edge's 32 levels on 16 lanes, enough for
    # every stencil row to be loaded once per lane block and shared
    "sw_fct_32levels": (3, 32, 2, 2, 1e5, 1e3, False, 600.0, 2),
    # levels on lanes with several tracer lanes per cell: the cells and
    # tracer kernels share an entity's rows across 16 (or 8) lanes, of which
    # only the lead lane of each level forms ke
    "sw_fct4_8levels": (3, 8, 4, 2, 1e5, 1e3, False, 600.0, 3),
    "sw_centred4_4levels": (3, 4, 4, 1, 1e5, 1e3, False, 600.0, 3),
    "sw_fct8_4levels": (3, 4, 8, 2, 1e5, 1e3, False, 600.0, 2),
    # the plain CPU-parity run: level 4, dt 960 s, viscosity 1e5, 100 steps
    "plain_L4_100": (4, 1, 0, 0, 1e5, 0.0, False, 960.0, 100),
    # the single-instance workload's configuration on a level-4 mesh, run
    # for the 1,000 steps the accuracy bound (1e-9 after 1,000 steps) names
    "workload_L4_1000": (4, 1, 4, 2, 1e5, 1e3, False, 240.0, 1000),
}


def _run_pair(name, steps=None, seed=1):
    from mpas_hip import Context, sw_workload_state
    from oracle.oracle import OracleSW
    lev, nL, nT, scheme, visc2, kappa2, freeze, dt, nst = FUSED_CASES[name]
    steps = nst if steps is None else steps
    m = mesh(lev)
    u, h, tr = sw_workload_state(m, nT, seed=seed)
    if nL > 1:  # independent levels: scaled copies
        h = np.stack([h * (1 + 0.01 * k) for k in range(nL)], axis=1)
        u = np.stack([u * (1 - 0.1 * k) for k in range(nL)], axis=1)
        tr = np.stack([tr] * nL, axis=1)
    h_s = np.full(m.nCells, 4500.0)  # apps/sw/initialize.cc:19-24
    ctx = Context(m, nLevels=nL, nTracers=nT, tracer_scheme=scheme, freeze_u=freeze, visc2=visc2,
                  kappa2=kappa2)
    ctx.sw_upload(h, u, tr if nT else None, h_s)
    ctx.rk4_step(dt, steps)
    got = ctx.sw_download()
    om = _omesh(m, nL=nL)
    o = OracleSW(om, nL=nL, nT=nT, tracer_adv=scheme, freeze_u=freeze, visc2=visc2,
                 kappa2=kappa2)
    o.set("h", h)
    o.set("u", u)
    if nT:
        o.set("tracers", tr)
    o.set("h_s", h_s)
    o.init_diagnostics()
    o.run(steps, dt)
    ref = (o.get("h"), o.get("u"), o.get("tracers") if nT else np.zeros(0))
    diag = {k: (ctx.sw_field(k), o.get(k)) for k in ("h_edge", "ke", "pv_edge", "vorticity",
                                                    "pv_vertex")}
    ctx.close()
    return got, ref, diag


@pytest.mark.parametrize("name", [k for k in FUSED_CASES if not k.endswith("_1000")])
def test_fused_step_matches_oracle(name):
    got, ref, diag = _run_pair(name)
    for g, r, nm in zip(got, ref, ("h", "u", "tracers")):
        assert max_rel(g, r) <= TOL, (nm, max_rel(g, r))
        assert bitwise_equal(g, r), (nm, max_rel(g, r))
    for k, (g, r) in diag.items():
        assert bitwise_equal(g, r), (k, max_rel(g, r))


def test_fused_1000_steps_within_1e9():
    """1,000 fused steps of the level-8 workload's configuration (4 FCT
    tracers, del2 mixing and viscosity) on level 4: within 1e-9 relative of
    the oracle, and in fact bit for bit."""
    got, ref, _ = _run_pair("workload_L4_1000")
    for g, r, nm in zip(got, ref, ("h", "u", "tracers")):
        print(nm, "max rel", max_rel(g, r))
        assert max_rel(g, r) <= 1e-9, nm
        assert bitwise_equal(g, r), nm


def test_fused_partial_levels_centred_mixing_bitwise(mesh3):
    """Several levels with partial columns (maxLevelCell < nLevels): the
    fused step (one level per lane) with centred tracers and del2 mixing,
    which reads maxLevelEdgeTop, against the oracle bit for bit."""
    from mpas_hip import TRACER_CENTRED, Context, sw_workload_state
    from oracle.oracle import OracleSW
    m, nL, nT = mesh3, 3, 2
    rng = np.random.default_rng(23)
    mlc = rng.integers(1, nL + 1, size=m.nCells).astype(np.int32)
    u, h, tr = sw_workload_state(m, nT, seed=4)
    h = np.stack([h * (1 + 0.01 * k) for k in range(nL)], axis=1)
    u = np.stack([u * (1 - 0.1 * k) for k in range(nL)], axis=1)
    tr = np.stack([tr * (1 + 0.2 * k) for k in range(nL)], axis=1)
    h_s = np.full(m.nCells, 4500.0)
    ctx = Context(m, nLevels=nL, nTracers=nT, tracer_scheme=TRACER_CENTRED, visc2=1e5,
                  kappa2=1e3, maxLevelCell=mlc)
    ctx.sw_upload(h, u, tr, h_s)
    ctx.rk4_step(600.0, 4)
    got = ctx.sw_download()
    o = OracleSW(_omesh(m, nL=nL, maxLevelCell=mlc), nL=nL, nT=nT, tracer_adv=TRACER_CENTRED,
                 visc2=1e5, kappa2=1e3)
    o.set("h", h)
    o.set("u", u)
    o.set("tracers", tr)
    o.set("h_s", h_s)
    o.init_diagnostics()
    o.run(4, 600.0)
    for g, r, nm in zip(got, (o.get("h"), o.get("u"), o.get("tracers")), ("h", "u", "tr")):
        assert bitwise_equal(g, r), (nm, max_rel(g, r))
    ctx.close()


@pytest.mark.parametrize("nL", [1, 2])
def test_fused_fct_boundary_mask_bitwise(mesh3, nL):
    """boundaryCell switches the third-order flux off next to boundaries
    (highOrderAdvMask, init_functions.cc:330-447): the fused FCT step with a
    random mask, one and two levels, against the oracle bit for bit."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    from oracle.oracle import OracleSW
    m, nT = mesh3, 2
    rng = np.random.default_rng(31 + nL)
    bnd = (rng.random(m.nCells * nL) < 0.15).astype(np.int32)
    u, h, tr = sw_workload_state(m, nT, seed=6)
    if nL > 1:
        h = np.stack([h * (1 + 0.01 * k) for k in range(nL)], axis=1)
        u = np.stack([u * (1 - 0.1 * k) for k in range(nL)], axis=1)
        tr = np.stack([tr] * nL, axis=1)
    h_s = np.full(m.nCells, 4500.0)
    ctx = Context(m, nLevels=nL, nTracers=nT, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3,
                  boundaryCell=bnd)
    ctx.sw_upload(h, u, tr, h_s)
    ctx.rk4_step(600.0, 3)
    got = ctx.sw_download()
    o = OracleSW(_omesh(m, nL=nL, boundaryCell=bnd), nL=nL, nT=nT, tracer_adv=TRACER_FCT,
                 visc2=1e5, kappa2=1e3)
    o.set("h", h)
    o.set("u", u)
    o.set("tracers", tr)
    o.set("h_s", h_s)
    o.init_diagnostics()
    o.run(3, 600.0)
    for g, r, nm in zip(got, (o.get("h"), o.get("u"), o.get("tracers")), ("h", "u", "tr")):
        assert bitwise_equal(g, r), (nm, max_rel(g, r))
    ctx.close()


def test_fused_wide_kernels_bitwise(monkeypatch):
    """The 8 / 16-wide stage kernels (cells with more than 6 edges) give the
    same bits as the oracle on the same workloads."""
    monkeypatch.setenv("MPAS_HIP_WIDE_KERNELS", "1")
    for name in ("sw_fct4", "sw_centred2", "sw_fct_32levels"):
        got, ref, diag = _run_pair(name, steps=2, seed=7)
        for g, r in zip(got, ref):
            assert bitwise_equal(g, r), name


def test_fused_one_step_each_config_bitwise():
    for name in FUSED_CASES:
        if name.endswith("_1000"):
            continue
        got, ref, _ = _run_pair(name, steps=1, seed=5)
        for g, r in zip(got, ref):
            assert bitwise_equal(g, r), name


def test_fused_deterministic(mesh3):
    a, _, _ = _run_pair("sw_fct4", steps=3, seed=9)
    b, _, _ = _run_pair("sw_fct4", steps=3, seed=9)
    for x, y in zip(a, b):
        assert bitwise_equal(x, y)


def test_mass_conserved_fused(mesh4):
    """Total mass sum(h * areaCell) is conserved to 1e-13 relative."""
    from mpas_hip import Context, TRACER_FCT, sw_workload_state
    m = mesh4
    u, h, tr = sw_workload_state(m, 2, seed=3)
    ctx = Context(m, nTracers=2, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    ctx.sw_upload(h, u, tr, np.full(m.nCells, 4500.0))
    area = m["areaCell"]
    m0 = np.sum(h * area)
    t0 = np.sum(tr * (h * area)[:, None], axis=0)
    ctx.rk4_step(960.0, 50)
    h1, _, tr1 = ctx.sw_download()
    assert abs(np.sum(h1 * area) / m0 - 1) < 1e-13
    t1 = np.sum(tr1.reshape(-1, 2) * (h1 * area)[:, None], axis=0)
    assert np.all(np.abs(t1 / t0 - 1) < 1e-12)
    ctx.close()


def test_error_paths_on_device(mesh3):
    from mpas_hip import Context, MpasError
    ctx = Context(mesh3)
    with pytest.raises(MpasError) as e:
        ctx.rk4_step(60.0, 1)  # no state uploaded yet
    assert e.value.code == -4
    with pytest.raises(MpasError) as e:
        ctx.sw_field("no_such_field")
    assert e.value.code in (-1, -4)
    ctx.close()


def test_compute_error_case1(mesh3):
    """compute_error (apps/sw/tasks/solve.cc:404-436) on device vs the oracle."""
    from mpas_hip import Context, sw_workload_state
    from oracle.oracle import OracleSW
    m = mesh3
    u, h = m.sw_case1()
    ctx = Context(m, freeze_u=True)
    dh = ctx.to_device(h)
    lat, lon = ctx.to_device(m["latCell"]), ctx.to_device(m["lonCell"])
    err = ctx.compute_error(dh, lat, lon, 0.0)
    o = OracleSW(_omesh(m), freeze_u=True)
    o.set("h", h)
    assert abs(err - o.compute_error()) <= 1e-12 * max(1.0, abs(err))
    ctx.close()


@pytest.mark.parametrize("order", [1, 2])
def test_reordered_context_bitwise(order):
    """A context on the mesh renumbered (reverse Cuthill-McKee or Morton
    cells; edges and vertices after them) advances the same state, seen
    through the permutation, to the same bits: every list keeps its order
    and the ascending-edge sums keep the original ids' order."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(4)
    u, h, tr = sw_workload_state(m, 4, seed=12)
    tr = tr.reshape(m.nCells, 4)
    h_s = np.full(m.nCells, 4500.0)
    ctx = Context(m, nTracers=4, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    r = ctx.reordered(order)
    cof, eof, _ = r.order()
    ctx.sw_upload(h, u, tr.ravel(), h_s)
    ctx.rk4_step(480.0, 4)
    ref = ctx.sw_download()
    r.sw_upload(h[cof], u[eof], tr[cof].ravel(), h_s[cof])
    r.rk4_step(480.0, 4)
    gh, gu, gt = r.sw_download()
    assert bitwise_equal(gh, ref[0][cof])
    assert bitwise_equal(gu, ref[1][eof])
    assert bitwise_equal(gt.reshape(-1, 4), ref[2].reshape(-1, 4)[cof])
    r.close()
    ctx.close()



@pytest.mark.gpu
@pytest.mark.parametrize("nT", [2, 4])
def test_cells_lds_staging_bitwise(monkeypatch, nT):
    """The one-level FCT cells kernel with its workgroup's edge values
    staged in LDS (MPAS_HIP_CELLS_LDS=1, read when the context is made)
    gives the bits of the default kernel, which the oracle tests pin; 2562
    cells leave the last workgroup partly empty."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(4)
    u, h, tr = sw_workload_state(m, nT, seed=21)
    h_s = np.full(m.nCells, 4500.0)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MPAS_HIP_CELLS_LDS", flag)
        ctx = Context(m, nTracers=nT, tracer_scheme=TRACER_FCT, visc2=1e5)
        ctx.sw_upload(h, u, tr, h_s)
        ctx.rk4_step(480.0, 3)
        out.append(ctx.sw_download())
        ctx.close()
    for a, b in zip(out[0], out[1]):
        assert bitwise_equal(a, b)


@pytest.mark.parametrize("var,nT", [("MPAS_HIP_EDGES_LDS", 0), ("MPAS_HIP_EDGES_LDS", 4),
                                    ("MPAS_HIP_HO_LDS", 2), ("MPAS_HIP_HO_LDS", 4),
                                    ("MPAS_HIP_HO_LDS", 8), ("MPAS_HIP_TRACER_LDS", 2),
                                    ("MPAS_HIP_TRACER_LDS", 4), ("MPAS_HIP_TRACER_LDS", 8),
                                    ("MPAS_HIP_CELLS_CST", 2), ("MPAS_HIP_CELLS_CST", 4),
                                    ("MPAS_HIP_CELLS_CST", 8)])
def test_lds_staging_bitwise(monkeypatch, var, nT):
    """The one-level momentum, high-order flux and tracer-update kernels
    with the values their stencils gather staged in LDS per workgroup (the
    default; the variable set to 0, read when the context is made, turns it
    off) give the bits of the global-gather kernels: 7680 edges make 30
    workgroups (60 for the flux kernel at 8 tracers, two lanes per edge);
    the tracer update's 2562 cells make 21, 41 and 81 workgroups of 128, 64
    and 32 cells at 2, 4 and 8 tracers, the last one partly empty, with
    del2 mixing on (its neighbours' tracers come from LDS too); the cells
    kernel (two tracers per lane) stages its cells' and their neighbours'
    tracers and thickness in 11, 21 and 41 workgroups of 256, 128 and 64
    cells and forms h_edge from the thickness."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(4)
    u, h, tr = sw_workload_state(m, nT, seed=22)
    h_s = np.full(m.nCells, 4500.0)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv(var, flag)
        ctx = Context(m, nTracers=nT, tracer_scheme=TRACER_FCT if nT else 0, visc2=1e5,
                      kappa2=1e3 if var == "MPAS_HIP_TRACER_LDS" else 0.0)
        ctx.sw_upload(h, u, tr if nT else None, h_s)
        ctx.rk4_step(480.0, 3)
        out.append(ctx.sw_download())
        ctx.close()
    for a, b in zip(out[0], out[1]):
        assert bitwise_equal(a, b)
