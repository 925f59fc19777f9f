"""Tangent-linear model (TLM) and discrete adjoint of the fused RK4 step
(FCT limiter factors frozen at the forward pass), through the C ABI.

How they are pinned (the reference has no derivative code):
  * dot-product test: <M dx, y> = <dx, M^T y> for random dx, y, over several
    checkpoint segments (tolerance 1e-10 relative);
  * the TLM against central finite differences of the frozen-limiter
    forward (the map it differentiates) and, with no tracers, of the plain
    forward;
  * the frozen-limiter replay at the unperturbed state reproduces the
    forward bit for bit (the tape holds the forward's own limiter factors);
  * the adjoint through a torch.autograd.Function wrapper (used only here):
    the gradient of an observation loss against the TLM's directional
    derivative of the same loss.
"""
import numpy as np
import pytest

from conftest import bitwise_equal, mesh

pytestmark = pytest.mark.gpu

DOT_TOL = 1e-10


def _setup(level, nT, K, visc2=1e5, kappa2=1e3, seed=1):
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(level)
    ctx = Context(m, nTracers=nT, tracer_scheme=TRACER_FCT if nT else 0, visc2=visc2,
                  kappa2=kappa2)
    u, h, tr = sw_workload_state(m, nT, seed=seed)
    ctx.sw_upload(h, u, tr.ravel() if nT else None, np.full(m.nCells, 4500.0))
    ctx.adj_init(K)
    x0 = ctx.state_arrays()
    x0[0].upload(u)
    x0[1].upload(h)
    if nT:
        x0[2].upload(tr.ravel())
    return m, ctx, x0, (u, h, tr.ravel())


def _rand_state(ctx, rng, scale=(1.0, 1.0, 0.01)):
    L = ctx.nLevels
    n = (ctx.nEdges * L, ctx.nCells * L, ctx.nCells * L * ctx.nTracers)
    host = [scale[i] * rng.standard_normal(n[i]) for i in range(3)]
    dev = ctx.state_arrays()
    for i in range(3):
        if n[i]:
            dev[i].upload(host[i])
    return host, dev


def _get(ctx, x):
    out = [x[0].numpy(), x[1].numpy()]
    out.append(x[2].numpy() if ctx.nTracers else np.zeros(0))
    return out


def _dot(a, b):
    return sum(float(np.dot(x, y)) for x, y in zip(a, b))


@pytest.mark.parametrize("nT,nsteps,K", [(0, 7, 3), (1, 10, 4), (2, 6, 6)])
def test_dot_product(nT, nsteps, K):
    m, ctx, x0, _ = _setup(4, nT, K)
    rng = np.random.default_rng(7 + nT)
    dt = 600.0
    dx_h, dx = _rand_state(ctx, rng)
    y_h, y = _rand_state(ctx, rng)
    dxN = ctx.state_arrays()
    ctx.tlm_forward(dt, nsteps, x0, dx, dx1=dxN)
    ctx.adj_forward(dt, nsteps, x0)
    xb = ctx.state_arrays()
    ctx.adj_backward(y, xb)
    ctx.sync()
    lhs = _dot(_get(ctx, dxN), y_h)
    rhs = _dot(dx_h, _get(ctx, xb))
    rel = abs(lhs - rhs) / abs(lhs)
    print(f"dot test nT={nT} steps={nsteps} K={K}: <M dx, y> = {lhs:.16e}, "
          f"<dx, M^T y> = {rhs:.16e}, rel {rel:.2e}")
    assert rel < DOT_TOL
    ctx.close()


def test_tlm_final_state_equals_forward():
    """The TLM's own forward trajectory is the fused forward's, bit for bit."""
    m, ctx, x0, _ = _setup(3, 2, 3)
    rng = np.random.default_rng(3)
    _, dx = _rand_state(ctx, rng)
    x1t, x1f = ctx.state_arrays(), ctx.state_arrays()
    ctx.tlm_forward(600.0, 5, x0, dx, x1=x1t)
    ctx.adj_forward(600.0, 5, x0, x1f)
    ctx.sync()
    for a, b in zip(_get(ctx, x1t), _get(ctx, x1f)):
        assert bitwise_equal(a, b)
    ctx.close()


def _fd_check(level, nT, nsteps, eps, frozen):
    """Central differences against the TLM.  With tracers the forward has
    kinks where F = u h_edge or an antidiffusive flux crosses zero (the
    upwind side, the 3rd-order sign, the limiter pair); the perturbation is
    made relative to the state (u and tracers scaled by themselves) so that
    no branch is crossed and the map is smooth over [x - eps dx, x + eps dx]."""
    m, ctx, x0, host0 = _setup(level, nT, K=nsteps)
    dt = 600.0
    rng = np.random.default_rng(11)
    dx_h, dx = _rand_state(ctx, rng)
    if nT:
        dx_h = [0.01 * host0[0] * dx_h[0], dx_h[1], host0[2] * dx_h[2]]
        for i in range(3):
            dx[i].upload(dx_h[i])
    dxN, xN = ctx.state_arrays(), ctx.state_arrays()
    ctx.tlm_forward(dt, nsteps, x0, dx, x1=xN, dx1=dxN)
    ctx.adj_forward(dt, nsteps, x0)
    if frozen:  # records the tape (segment 0 = the whole run)
        yb, xb = ctx.state_arrays(), ctx.state_arrays()
        ctx.adj_backward(yb, xb)
    xp, xm, out_p, out_m = (ctx.state_arrays() for _ in range(4))
    for i in range(3):
        if dx_h[i].size:
            xp[i].upload(host0[i] + eps * dx_h[i])
            xm[i].upload(host0[i] - eps * dx_h[i])
    if frozen:
        ctx.adj_frozen_forward(nsteps, xp, out_p)
        ctx.adj_frozen_forward(nsteps, xm, out_m)
        same = ctx.state_arrays()
        ctx.adj_frozen_forward(nsteps, x0, same)
        ctx.sync()
        for a, b in zip(_get(ctx, same), _get(ctx, xN)):
            assert bitwise_equal(a, b), "frozen replay differs from the forward"
    else:
        ctx.adj_forward(dt, nsteps, xp, out_p)
        ctx.adj_forward(dt, nsteps, xm, out_m)
    ctx.sync()
    tl = _get(ctx, dxN)
    p, q = _get(ctx, out_p), _get(ctx, out_m)
    rels = []
    for name, a, b, t in zip(("u", "h", "tracers"), p, q, tl):
        if not t.size:
            continue
        fd = (a - b) / (2 * eps)
        rel = np.linalg.norm(fd - t) / np.linalg.norm(t)
        print(f"FD vs TLM {name}: |TLM| {np.linalg.norm(t):.3e}, rel diff {rel:.2e}")
        rels.append(rel)
    ctx.close()
    return rels


def test_tlm_vs_finite_differences_sw():
    # central differences: ~2e-10 at eps = 1e-4 (truncation and round-off)
    assert max(_fd_check(4, 0, 6, 1e-4, frozen=False)) < 1e-6


@pytest.mark.parametrize("eps", [1e-4, 1e-6])
def test_tlm_vs_finite_differences_frozen_fct(eps):
    # ~2e-9 at eps = 1e-4, ~2e-7 at 1e-6 (round-off of the differences)
    assert max(_fd_check(4, 2, 6, eps, frozen=True)) < 1e-5


def test_torch_autograd_observation_loss():
    """The torch.autograd.Function wrapper (torch_adjoint_check.py) in a
    process of its own: torch ships its own HIP runtime under the same
    soname as ROCm's, so torch must load before the library does."""
    import json
    import os
    import subprocess
    import sys
    pytest.importorskip("torch")
    script = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torch_adjoint_check.py")
    r = subprocess.run([sys.executable, script], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    res = json.loads(r.stdout.strip().splitlines()[-1])
    assert res["grad_finite"] and res["rel_adjoint_vs_tlm"] < DOT_TOL


def test_tile_and_schedule_keep_the_gradient():
    """Workgroup sizes of the adjoint and TLM kernels, the replay beside the
    sweep (second tape and stream) or in turn, and graph-replayed segment
    sweeps: the gradient and the TLM are the same bits in every
    configuration and after mpas_adj_autotune, over several segments."""
    m, ctx, x0, _ = _setup(4, 2, K=3)
    rng = np.random.default_rng(21)
    y_h, y = _rand_state(ctx, rng)
    dx_h, dx = _rand_state(ctx, rng)
    dt, nsteps = 600.0, 8  # segments of 3, 3 and 2 steps

    def run():
        xb, dxN = ctx.state_arrays(), ctx.state_arrays()
        ctx.adj_forward(dt, nsteps, x0)
        ctx.adj_backward(y, xb)
        ctx.adj_backward(y, ctx.state_arrays())  # a second sweep replays the graphs
        ctx.tlm_forward(dt, nsteps, x0, dx, dx1=dxN)
        ctx.sync()
        return _get(ctx, xb) + _get(ctx, dxN)

    ref = run()
    for ad, tl, ov, gr in ((64, 512, True, True), (512, 64, False, True), (128, 128, True, False),
                           (256, 256, False, False)):
        ctx.adj_config(ad, tl, ov, gr)
        assert ctx.adj_get_config() == dict(ad_block=ad, tl_block=tl, overlap=int(ov),
                                            graph=int(gr))
        for a, b in zip(run(), ref):
            assert bitwise_equal(a, b), (ad, tl, ov, gr)
    trials = ctx.adj_autotune(dt, nsteps)
    assert len(trials) >= 8 and all(t["ms"] > 0 for t in trials)
    print("adjoint tuned:", ctx.adj_get_config())
    for a, b in zip(run(), ref):
        assert bitwise_equal(a, b), "tuned"
    ctx.close()


@pytest.mark.parametrize("nL", [1, 8])
def test_fused_prep_keeps_the_gradient(monkeypatch, nL):
    """The stage-0 kernels of a step's sweep running the next step's prep
    (the default) give the bits of the separate prep kernel
    (MPAS_HIP_ADJ_FUSE_PREP=0, read at mpas_adj_init), over segments of
    several steps, with one level and with levels on lanes."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(4)
    u, h, tr = sw_workload_state(m, 2, seed=5)
    rng = np.random.default_rng(33)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MPAS_HIP_ADJ_FUSE_PREP", flag)
        ctx = Context(m, nLevels=nL, nTracers=2, tracer_scheme=TRACER_FCT, visc2=1e5,
                      kappa2=1e3)
        rep = lambda a: np.repeat(a.reshape(a.shape[0], 1, -1), nL, axis=1).ravel()
        ctx.sw_upload(rep(h.reshape(-1, 1)), rep(u.reshape(-1, 1)), rep(tr.reshape(m.nCells, -1)),
                      np.full(m.nCells, 4500.0))
        ctx.adj_init(3)
        x0 = ctx.state_arrays()
        x0[0].upload(rep(u.reshape(-1, 1)))
        x0[1].upload(rep(h.reshape(-1, 1)))
        x0[2].upload(rep(tr.reshape(m.nCells, -1)))
        y_h, y = _rand_state(ctx, np.random.default_rng(33))
        xb = ctx.state_arrays()
        ctx.adj_forward(600.0, 7, x0)  # segments of 3, 3 and 1 steps
        ctx.adj_backward(y, xb)
        ctx.sync()
        out.append(_get(ctx, xb))
        ctx.close()
    for a, b in zip(out[0], out[1]):
        assert bitwise_equal(a, b)


@pytest.mark.parametrize("nL", [1, 2])
def test_merged_vertex_part_keeps_the_gradient(monkeypatch, nL):
    """Below 8 levels the vertex part runs in the cell part's launch, whose
    cells form their vertices' h_vertex adjoints themselves: the gradient
    and the TLM are the bits of the two separate launches
    (MPAS_HIP_ADJ_MERGE_VERTEX=0, read at mpas_adj_init), over segments of
    several steps."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(4)
    u, h, tr = sw_workload_state(m, 2, seed=9)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MPAS_HIP_ADJ_MERGE_VERTEX", flag)
        ctx = Context(m, nLevels=nL, nTracers=2, tracer_scheme=TRACER_FCT, visc2=1e5,
                      kappa2=1e3)
        rep = lambda a: np.repeat(a.reshape(a.shape[0], 1, -1), nL, axis=1).ravel()
        ctx.sw_upload(rep(h.reshape(-1, 1)), rep(u.reshape(-1, 1)), rep(tr.reshape(m.nCells, -1)),
                      np.full(m.nCells, 4500.0))
        ctx.adj_init(3)
        x0 = ctx.state_arrays()
        x0[0].upload(rep(u.reshape(-1, 1)))
        x0[1].upload(rep(h.reshape(-1, 1)))
        x0[2].upload(rep(tr.reshape(m.nCells, -1)))
        y_h, y = _rand_state(ctx, np.random.default_rng(43))
        xb = ctx.state_arrays()
        ctx.adj_forward(600.0, 7, x0)  # segments of 3, 3 and 1 steps
        ctx.adj_backward(y, xb)
        ctx.adj_backward(y, ctx.state_arrays())  # a second sweep replays the graphs
        ctx.sync()
        out.append(_get(ctx, xb))
        ctx.close()
    for a, b in zip(out[0], out[1]):
        assert bitwise_equal(a, b)


@pytest.mark.parametrize("nL", [1, 8])
def test_paired_stencils_keep_the_gradient(monkeypatch, nL):
    """The edge part reading the slot-paired stencils, with the transposed
    edgesOnEdge weights found through each slot's reverse slot (the default
    on meshes of at most hexagons), against the slot-major rows and the
    transposed lists (MPAS_HIP_ADJ_PAIRED=0, read at mpas_adj_init): the
    sums run in another order, so the gradients agree to rounding
    (1e-12 relative to the largest component) and both pass the dot-product
    test, with one level and with levels on lanes."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(4)
    u, h, tr = sw_workload_state(m, 2, seed=7)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MPAS_HIP_ADJ_PAIRED", flag)
        ctx = Context(m, nLevels=nL, nTracers=2, tracer_scheme=TRACER_FCT, visc2=1e5,
                      kappa2=1e3)
        rep = lambda a: np.repeat(a.reshape(a.shape[0], 1, -1), nL, axis=1).ravel()
        ctx.sw_upload(rep(h.reshape(-1, 1)), rep(u.reshape(-1, 1)), rep(tr.reshape(m.nCells, -1)),
                      np.full(m.nCells, 4500.0))
        ctx.adj_init(3)
        x0 = ctx.state_arrays()
        x0[0].upload(rep(u.reshape(-1, 1)))
        x0[1].upload(rep(h.reshape(-1, 1)))
        x0[2].upload(rep(tr.reshape(m.nCells, -1)))
        rng = np.random.default_rng(41)
        y_h, y = _rand_state(ctx, rng)
        dx_h, dx = _rand_state(ctx, rng)
        xb, dxN = ctx.state_arrays(), ctx.state_arrays()
        ctx.adj_forward(600.0, 7, x0)  # segments of 3, 3 and 1 steps
        ctx.adj_backward(y, xb)
        ctx.tlm_forward(600.0, 7, x0, dx, dx1=dxN)
        ctx.sync()
        g = _get(ctx, xb)
        lhs, rhs = _dot(_get(ctx, dxN), y_h), _dot(g, dx_h)
        assert abs(lhs - rhs) <= DOT_TOL * abs(lhs), (flag, lhs, rhs)
        out.append(g)
        ctx.close()
    for a, b in zip(out[0], out[1]):
        scale = np.max(np.abs(a))
        assert np.max(np.abs(a - b)) <= 1e-12 * scale


def test_uniform_entity_path_keeps_the_gradient(monkeypatch):
    """A 64-level context's edge and cell parts take their entity from the
    wave's first lane (the default; MPAS_HIP_ADJ_UNIFORM=0, read at
    mpas_adj_init, keeps per-lane indices): the same gradient bits."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(3)
    nM, nT = 64, 2
    members = [sw_workload_state(m, nT, seed=60 + g, u0=10.0 + 0.2 * g) for g in range(nM)]
    U, H, T = _pack_members(m, [(u, h, tr.ravel()) for u, h, tr in members], nT)
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MPAS_HIP_ADJ_UNIFORM", flag)
        ctx = Context(m, nLevels=nM, nTracers=nT, tracer_scheme=TRACER_FCT, visc2=1e5,
                      kappa2=1e3)
        ctx.sw_upload(H, U, T, np.full(m.nCells, 4500.0))
        ctx.adj_init(3)
        x0 = ctx.state_arrays()
        for a, v in zip(x0, (U, H, T)):
            a.upload(v)
        y_h, y = _rand_state(ctx, np.random.default_rng(47))
        xb = ctx.state_arrays()
        ctx.adj_forward(600.0, 7, x0)  # segments of 3, 3 and 1 steps
        ctx.adj_backward(y, xb)
        ctx.sync()
        out.append(_get(ctx, xb))
        ctx.close()
    for a, b in zip(out[0], out[1]):
        assert bitwise_equal(a, b)


def test_adjoint_error_paths():
    from mpas_hip import TRACER_CENTRED, Context
    from mpas_hip._lib import MpasError
    m = mesh(3)
    ctx = Context(m, nTracers=1, tracer_scheme=TRACER_CENTRED)
    with pytest.raises(MpasError):
        ctx.adj_init(4)  # the adjoint covers FCT tracers only
    ctx.close()
    m, ctx, x0, _ = _setup(3, 1, K=2)
    xb = ctx.state_arrays()
    with pytest.raises(MpasError):
        ctx.adj_backward(x0, xb)  # no forward yet
    ctx.adj_forward(600.0, 5, x0)
    ctx.adj_backward(x0, xb)  # 3 segments: the tape ends on segment 0 (2 steps)
    with pytest.raises(MpasError):
        ctx.adj_frozen_forward(3, x0, xb)
    with pytest.raises(MpasError):
        ctx.adj_init(0)
    ctx.close()


def _pack_members(m, members, nT):
    """[entity][member] (u, h) and [cell][member][tracer] tracers."""
    U = np.stack([x[0] for x in members], axis=1).ravel()
    H = np.stack([x[1] for x in members], axis=1).ravel()
    T = np.stack([x[2].reshape(m.nCells, nT) for x in members], axis=1).ravel()
    return U, H, T


@pytest.mark.parametrize("nM", [6, 12, 64])
def test_members_as_levels_equal_single_runs(nM):
    """Levels of one context are independent, so a context whose levels are
    ensemble members (the batched adjoint) gives every member the final
    state and gradient of its own single-level run, bit for bit; the
    levels run one per lane (adjacent lanes, adjacent members).  With 12
    members the adjoint's edge and cell parts take one lane per entity and
    member, which adds the partial sums the single run spreads over 2 or 4
    lanes in the same tree; with 64 each wave holds one entity, whose static
    rows those parts load as scalars."""
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(3)
    nT, K, nsteps, dt = 2, 3, 7, 600.0
    kw = dict(nTracers=nT, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    members = []
    for g in range(nM):
        u, h, tr = sw_workload_state(m, nT, seed=20 + g, u0=12.0 + 2 * g, bump_amp=6.0 + g)
        members.append((u, h, tr.ravel()))
    U, H, T = _pack_members(m, members, nT)
    cm = Context(m, nLevels=nM, **kw)
    cm.sw_upload(H, U, T, np.full(m.nCells, 4500.0))
    cm.adj_init(K)
    x0 = cm.state_arrays()
    for a, v in zip(x0, (U, H, T)):
        a.upload(v)
    x1 = cm.state_arrays()
    cm.adj_forward(dt, nsteps, x0, x1)
    rng = np.random.default_rng(11)
    E, C = m.nEdges, m.nCells
    yb = [rng.standard_normal(E * nM), rng.standard_normal(C * nM),
          0.01 * rng.standard_normal(C * nM * nT)]
    y = cm.state_arrays()
    for a, v in zip(y, yb):
        a.upload(v)
    xb = cm.state_arrays()
    cm.adj_backward(y, xb)
    cm.sync()
    f = [x1[0].numpy().reshape(E, nM), x1[1].numpy().reshape(C, nM),
         x1[2].numpy().reshape(C, nM, nT)]
    g = [xb[0].numpy().reshape(E, nM), xb[1].numpy().reshape(C, nM),
         xb[2].numpy().reshape(C, nM, nT)]
    c1 = Context(m, **kw)
    c1.sw_upload(members[0][1], members[0][0], members[0][2], np.full(C, 4500.0))
    c1.adj_init(K)
    for k, (u, h, tr) in enumerate(members):
        a0 = c1.state_arrays()
        for a, v in zip(a0, (u, h, tr)):
            a.upload(v)
        a1 = c1.state_arrays()
        c1.adj_forward(dt, nsteps, a0, a1)
        yk = c1.state_arrays()
        yk[0].upload(yb[0].reshape(E, nM)[:, k].copy())
        yk[1].upload(yb[1].reshape(C, nM)[:, k].copy())
        yk[2].upload(yb[2].reshape(C, nM, nT)[:, k, :].ravel().copy())
        bk = c1.state_arrays()
        c1.adj_backward(yk, bk)
        c1.sync()
        assert bitwise_equal(a1[0].numpy(), f[0][:, k].copy()), k
        assert bitwise_equal(a1[1].numpy(), f[1][:, k].copy()), k
        assert bitwise_equal(a1[2].numpy(), f[2][:, k, :].ravel().copy()), k
        assert bitwise_equal(bk[0].numpy(), g[0][:, k].copy()), k
        assert bitwise_equal(bk[1].numpy(), g[1][:, k].copy()), k
        assert bitwise_equal(bk[2].numpy(), g[2][:, k, :].ravel().copy()), k
    # the dot-product test on the multi-level tape
    dx_h, dx = _rand_state(cm, rng)
    dxN = cm.state_arrays()
    cm.tlm_forward(dt, nsteps, x0, dx, dx1=dxN)
    cm.adj_forward(dt, nsteps, x0)
    xb2 = cm.state_arrays()
    cm.adj_backward(y, xb2)
    cm.sync()
    lhs = _dot(_get(cm, dxN), yb)
    rhs = _dot(dx_h, _get(cm, xb2))
    assert abs(lhs - rhs) / abs(lhs) < DOT_TOL
    c1.close()
    cm.close()
