"""Ensembles (mpas_batch_*): members on meshes of uneven size advanced by
one launch per phase for all member groups (direct or replayed from a
hipGraph), members of one mesh and time step stored member-minor.

Design claim: a member evolves bit for bit as it does alone (every kernel
treats levels, here the members, independently).  Pinned against a plain
context run of each member, across launch modes and configurations."""
import numpy as np
import pytest

from conftest import bitwise_equal, mesh

pytestmark = pytest.mark.gpu

LEVELS = (3, 4, 3, 4, 3, 3, 4)
DTS = (600.0, 300.0, 450.0, 300.0, 600.0, 600.0, 300.0)


def _ensemble(levels=LEVELS, nT=2, dts=DTS):
    from mpas_hip import TRACER_FCT, Batch, Context, sw_workload_state
    lv = sorted(set(levels))
    ctxs = [Context(mesh(L), nTracers=nT, tracer_scheme=TRACER_FCT if nT else 0, visc2=1e5,
                    kappa2=1e3 if nT else 0.0)
            for L in lv]
    mesh_of = [lv.index(L) for L in levels]
    states = []
    for i, L in enumerate(levels):
        u, h, tr = sw_workload_state(mesh(L), nT, seed=10 + i, u0=10.0 + 3 * i,
                                     bump_amp=5.0 + 2 * i)
        states.append((h, u, tr.ravel() if nT else None, np.full(mesh(L).nCells, 4500.0)))
    b = Batch(ctxs, mesh_of, dts)
    return ctxs, b, states, mesh_of


def _load(b, states):
    for i, (h, u, tr, hs) in enumerate(states):
        b.upload(i, h, u, tr, hs)


def _singles(ctxs, states, mesh_of, nsteps, dts=DTS):
    out = []
    for i, (h, u, tr, hs) in enumerate(states):
        c = ctxs[mesh_of[i]]
        c.sw_upload(h, u, tr, hs)
        c.rk4_step(dts[i], nsteps)
        out.append(c.sw_download())
    return out


def _close(ctxs, b):
    b.close()
    for c in ctxs:
        c.close()


def _oracle_member(level, nT, state, dt, nsteps):
    """One member advanced by the oracle (the CPU restatement of the
    reference's step) with the context settings of _ensemble."""
    from mpas_hip import TRACER_FCT
    from oracle.oracle import OracleMesh, OracleSW
    h, u, tr, hs = state
    o = OracleSW(OracleMesh(mesh(level), nL=1), nL=1, nT=nT,
                 tracer_adv=TRACER_FCT if nT else 0, visc2=1e5, kappa2=1e3 if nT else 0.0)
    o.set("h", h)
    o.set("u", u)
    if nT:
        o.set("tracers", tr)
    o.set("h_s", hs)
    o.init_diagnostics()
    o.run(nsteps, dt)
    return o.get("h"), o.get("u"), o.get("tracers") if nT else np.zeros(0)


@pytest.mark.parametrize("nT", [2, 4])
def test_members_match_oracle(nT):
    """Every batched member against the oracle's own run of that member,
    bit for bit (not only through a single-context run): three groups of
    different meshes and time steps, in two configurations (one member per
    lane with one entity per wave where it fills the wave; two members per
    lane, levels-on-lanes layouts only)."""
    from mpas_hip import BATCH_GRAPH
    ctxs, b, states, mesh_of = _ensemble(nT=nT)
    nsteps = 3
    ref = [_oracle_member(LEVELS[i], nT, states[i], DTS[i], nsteps) for i in range(len(LEVELS))]
    for lpl, we in ((1, 1), (2, 0)):
        b.config([256] * 6, lpl, 0, 1, we)
        _load(b, states)
        b.step(nsteps, BATCH_GRAPH)
        b.sync()
        for i in range(b.n):
            for a, r, nm in zip(b.download(i), ref[i], ("h", "u", "tracers")):
                assert bitwise_equal(a, r), (nT, lpl, we, i, nm)
    _close(ctxs, b)


def test_groups_follow_mesh_and_dt():
    ctxs, b, _, _ = _ensemble()
    # (L3, 600), (L4, 300), (L3, 450)
    assert b.groups() == 3
    _close(ctxs, b)


def test_members_equal_single_runs_both_modes():
    from mpas_hip import BATCH_GRAPH, BATCH_LAUNCH
    ctxs, b, states, mesh_of = _ensemble()
    ref = _singles(ctxs, states, mesh_of, 5)
    for mode in (BATCH_GRAPH, BATCH_LAUNCH):
        _load(b, states)
        b.step(5, mode)
        b.sync()
        for i in range(b.n):
            for a, r in zip(b.download(i), ref[i]):
                assert bitwise_equal(a, r), (mode, i)
    _close(ctxs, b)


@pytest.mark.parametrize("nT", [0, 2, 4])
def test_persistent_launch_keeps_results(nT):
    """Every phase of several steps in one cooperative launch (grid barriers
    between phases): every member equals its single run bit for bit, in
    several layouts and placements; the tuned mode is one of the two."""
    from mpas_hip import BATCH_GRAPH, BATCH_PERSISTENT, BATCH_TUNED
    levels = (3, 4, 3, 4, 4, 3, 4, 4, 5, 3, 5)  # groups of 4, 5 and 2 members
    dts = tuple({3: 600.0, 4: 300.0, 5: 150.0}[L] for L in levels)
    ctxs, b, states, mesh_of = _ensemble(levels, nT, dts)
    ref = _singles(ctxs, states, mesh_of, 23, dts)  # more than one launch of 20 steps
    for lpl, order, xcd, we in ((1, 0, 1, 1), (1, 1, 0, 0), (2, 0, 1, 2)):
        b.config([256] * 6, lpl, order, xcd, we)
        _load(b, states)
        print(f"persistent: nT {nT}, members/lane {lpl}, order {order}, xcd {xcd}, "
              f"entity/wave {we}", flush=True)
        b.step(23, BATCH_PERSISTENT)
        b.sync()
        for i in range(b.n):
            for a, r in zip(b.download(i), ref[i]):
                assert bitwise_equal(a, r), (nT, lpl, order, xcd, we, i)
    _load(b, states)
    b.autotune(trial_steps=2)
    assert b.tuned_mode() in (BATCH_GRAPH, BATCH_PERSISTENT)
    _load(b, states)
    b.step(23, BATCH_TUNED)
    b.sync()
    for i in range(b.n):
        for a, r in zip(b.download(i), ref[i]):
            assert bitwise_equal(a, r), ("tuned", i)
    _close(ctxs, b)


def test_mesh_orders_keep_results():
    """The groups on reordered meshes (reverse Cuthill-McKee, Morton, back to
    the given order): uploads and downloads keep the callers' numbering and
    every member equals its single run bit for bit, also across a change of
    order between steps; the autotuner's choice keeps the results."""
    from mpas_hip import BATCH_GRAPH, BATCH_PERSISTENT
    ctxs, b, states, mesh_of = _ensemble()
    ref = _singles(ctxs, states, mesh_of, 4)
    for order in (1, 2, 0):
        b.set_order(order)
        assert b.get_order() == order
        _load(b, states)
        b.step(2, BATCH_GRAPH)
        b.set_order((order + 1) % 3)  # the state follows the new numbering
        b.step(2, BATCH_PERSISTENT)
        b.sync()
        for i in range(b.n):
            for a, r in zip(b.download(i), ref[i]):
                assert bitwise_equal(a, r), (order, i)
    _load(b, states)
    trials = b.autotune(trial_steps=2)
    assert sum(t["trial"] == "mesh order" for t in trials) == 2
    b.step(4)
    b.sync()
    for i in range(b.n):
        for a, r in zip(b.download(i), ref[i]):
            assert bitwise_equal(a, r), ("tuned", i)
    _close(ctxs, b)


def test_configurations_keep_results():
    ctxs, b, states, mesh_of = _ensemble()
    ref = _singles(ctxs, states, mesh_of, 4)
    for blk, lpl, order, xcd, we in (([64] * 6, 1, 1, 0, 0), ([512] * 6, 2, 0, 1, 0),
                                     ([128, 256, 64, 512, 128, 64], 3, 1, 1, 1),
                                     ([256] * 6, 8, 0, 0, 2), ([256] * 6, 1, 0, 1, 2)):
        b.config(blk, lpl, order, xcd, we)
        _load(b, states)
        b.step(4)
        b.sync()
        for i in range(b.n):
            for a, r in zip(b.download(i), ref[i]):
                assert bitwise_equal(a, r), (blk, lpl, order, xcd, we, i)
    _close(ctxs, b)


@pytest.mark.parametrize("nT", [2, 3])
def test_member_pairs_keep_results(nT):
    """Groups with an even member count run the momentum, high-order flux and
    diagnostic kernels two members per lane (levels on lanes, or one entity
    per wave); odd groups and more members per lane do not."""
    levels = (3, 4, 3, 4, 4, 3, 4, 4, 5, 3, 5)  # groups of 4, 5 and 2 members
    dts = tuple({3: 600.0, 4: 300.0, 5: 150.0}[L] for L in levels)
    ctxs, b, states, mesh_of = _ensemble(levels, nT, dts)
    ref = _singles(ctxs, states, mesh_of, 3, dts)
    for lpl, we in ((1, 0), (1, 1), (1, 2), (2, 1)):
        b.config([256] * 6, lpl, 0, 1, we)
        _load(b, states)
        b.step(3)
        b.sync()
        for i in range(b.n):
            for a, r in zip(b.download(i), ref[i]):
                assert bitwise_equal(a, r), (nT, lpl, we, i)
    _close(ctxs, b)


def test_shared_stencil_rows_keep_results():
    """A group of 32 members: with two members per lane an edge's members
    fill 16 lanes and a cell's 32, so the kernels load each stencil row once
    per lane block and share it across lanes; every member still equals its
    single run."""
    levels = (3,) * 32 + (4,) * 3
    dts = tuple(600.0 if L == 3 else 300.0 for L in levels)
    ctxs, b, states, mesh_of = _ensemble(levels, 2, dts)
    ref = _singles(ctxs, states, mesh_of, 2, dts)
    for we in (0, 1):
        b.config([256] * 6, 1, 0, 1, we)
        _load(b, states)
        b.step(2)
        b.sync()
        for i in range(b.n):
            for a, r in zip(b.download(i), ref[i]):
                assert bitwise_equal(a, r), (we, i)
    _close(ctxs, b)


def test_tracer_lanes_share_rows_keep_results():
    """Four tracers: the cells and tracer-update kernels run two tracer lanes
    per member, and with levels on lanes a cell's lanes share its edge rows.
    A group of 4 members (8 lanes per cell) and one of 15 members, whose last
    level chunk is short with two members per lane (its lanes take fewer
    passes of the level loop than the others): every member equals its
    single run bit for bit."""
    levels = (3,) * 4 + (4,) * 15
    dts = tuple(600.0 if L == 3 else 300.0 for L in levels)
    ctxs, b, states, mesh_of = _ensemble(levels, 4, dts)
    ref = _singles(ctxs, states, mesh_of, 3, dts)
    for lpl, we in ((1, 0), (2, 0), (2, 1), (1, 2)):
        b.config([256] * 6, lpl, 0, 1, we)
        _load(b, states)
        b.step(3)
        b.sync()
        for i in range(b.n):
            for a, r in zip(b.download(i), ref[i]):
                assert bitwise_equal(a, r), (lpl, we, i)
    _close(ctxs, b)


def test_autotune_keeps_results_and_state():
    ctxs, b, states, mesh_of = _ensemble()
    ref = _singles(ctxs, states, mesh_of, 3)
    _load(b, states)
    before = [b.download(i) for i in range(b.n)]
    trials = b.autotune(trial_steps=2)
    assert len(trials) >= 5 and all(t["ms_per_step"] > 0 for t in trials)
    assert all(len(t["layout_block"]) == 36 for t in trials if "block" in t)
    print("tuned:", b.get_config())
    cfg, lb = b.get_config(), b.layout_blocks()
    # the tuned configuration is reproducible: config + layout blocks
    b.config(cfg["block"], cfg["members_per_lane"], cfg["order"], cfg["xcd"],
             cfg["wave_entity"])
    assert not b.layout_blocks().any()  # an explicit configuration clears them
    b.set_layout_blocks(lb)
    assert np.array_equal(b.layout_blocks(), lb)
    # a second run starts from the kernels' own sizes again
    trials2 = b.autotune(trial_steps=1)
    assert not any(trials2[0]["layout_block"])
    for x, y in zip(before, [b.download(i) for i in range(b.n)]):  # the trials are undone
        for a, c in zip(x, y):
            assert bitwise_equal(a, c)
    b.step(3)
    b.sync()
    for i in range(b.n):
        for a, r in zip(b.download(i), ref[i]):
            assert bitwise_equal(a, r)
    _close(ctxs, b)


def test_batch_error_paths():
    from mpas_hip import TRACER_CENTRED, TRACER_FCT, Batch, Context, MpasError
    a = Context(mesh(3), nTracers=2, tracer_scheme=TRACER_FCT)
    c = Context(mesh(3), nTracers=1, tracer_scheme=TRACER_FCT)
    with pytest.raises(MpasError):
        Batch([a, c], [0, 1], [600.0, 600.0])  # tracer counts differ
    with pytest.raises(MpasError):
        Batch([a], [1], [600.0])  # mesh index out of range
    b = Batch([a], [0, 0], [600.0, 600.0])
    with pytest.raises(MpasError):
        b.step(1)  # no member state yet
    with pytest.raises(MpasError):
        b.config([96] * 6)
    m = mesh(3)
    u = np.zeros(m.nEdges)
    h = np.full(m.nCells, 5000.0)
    tr = np.zeros(2 * m.nCells)
    b.upload(0, h, u, tr, np.full(m.nCells, 4500.0))
    with pytest.raises(MpasError):  # a group shares its bottom topography
        b.upload(1, h, u, tr, np.full(m.nCells, 4400.0))
    b.close()
    a.close()
    c.close()
    d = Context(mesh(3), nTracers=1, tracer_scheme=TRACER_CENTRED)
    b = Batch([d], [0], [600.0])  # centred tracers batch as well
    b.close()
    d.close()
    # 2,200 members of 8 tracers on level 5: the antidiffusive flux array
    # (edges x members x tracers) would pass the 4 GiB the kernels address
    big = Context(mesh(5), nTracers=8, tracer_scheme=TRACER_FCT)
    with pytest.raises(MpasError) as e:
        Batch([big], [0] * 2200, [600.0] * 2200)
    assert "4 GiB" in str(e.value)
    big.close()
