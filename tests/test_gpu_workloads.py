"""The named workloads at their full sizes, through the C ABI, each checked
against the oracle (the CPU restatement of the reference's tasks), the
committed golden fixture, or single runs, as its kind allows:

  * single instance: icosahedral level 8 (655,362 cells), FP64 RK4 with
    dt = 60 s, 4 FCT tracers, del2 mixing 1e3 and viscosity 1e5: two fused
    steps bit for bit against the oracle; 1,000 steps against the oracle's
    1,000-step fixture (tests/golden/sw_fct4_L8.json) within 1e-9 relative,
    and bit for bit, with mass conserved to 1e-13;
  * adjoint: level 6, 200 steps of dt = 240 s, 1 FCT tracer, checkpoints
    every 20 steps: the dot-product test over the whole run, the gradient of
    the 1,000-cell observation loss against central differences, and the
    tangent-linear model against central differences of the frozen-limiter
    forward;
  * ensemble: 96 x level 5 + 128 x level 6 + 32 x level 7 members, 2 FCT
    tracers, graph-replayed steps in the autotuned configuration: sampled
    members bit for bit against their single runs, one level-5 member
    against the oracle;
  * batched adjoint ensemble: 64 x level 7 + 64 x level 5 members, 100
    steps, checkpoints every 10 steps, 2,000 observation cells per member:
    sampled members' gradients bit for bit against their single runs, and
    the dot-product test on each multi-member tape.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import ROOT, bitwise_equal, max_rel, mesh

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(ROOT, "tests", "golden", "sw_fct4_L8.json")
FCT = dict(visc2=1e5, kappa2=1e3)


def _say(*a):
    print(*a, flush=True)


# ---- single instance, level 8 ------------------------------------------------

def _l8_state():
    from mpas_hip import sw_workload_state
    m = mesh(8)
    u, h, tr = sw_workload_state(m, 4, seed=1)
    return m, u, h, tr.ravel(), np.full(m.nCells, 4500.0)


def _l8_gpu(steps):
    from mpas_hip import TRACER_FCT, Context
    m, u, h, tr, h_s = _l8_state()
    ctx = Context(m, nTracers=4, tracer_scheme=TRACER_FCT, **FCT)
    ctx.sw_upload(h, u, tr, h_s)
    _say(f"level 8: {steps} fused steps on the device")
    ctx.rk4_step(60.0, steps)
    out = ctx.sw_download()
    ctx.close()
    return m, (u, h, tr), out


def test_level8_two_steps_bitwise_vs_oracle():
    from oracle.oracle import OracleMesh, OracleSW
    m, (u, h, tr), (gh, gu, gt) = _l8_gpu(2)
    _, _, _, _, h_s = _l8_state()
    _say("level 8: oracle setup and 2 steps (about 15 s)")
    o = OracleSW(OracleMesh(m), nT=4, tracer_adv=2, **FCT)
    o.set("h", h)
    o.set("u", u)
    o.set("tracers", tr)
    o.set("h_s", h_s)
    o.init_diagnostics()
    o.run(2, 60.0)
    for g, r, name in ((gh, o.get("h"), "h"), (gu, o.get("u"), "u"),
                       (gt, o.get("tracers"), "tracers")):
        assert max_rel(g, r) <= 1e-12, (name, max_rel(g, r))
        assert bitwise_equal(g, r), (name, max_rel(g, r))


def _sha(x):
    return hashlib.sha256(np.ascontiguousarray(x, dtype=np.float64).tobytes()).hexdigest()


def test_level8_1000_steps_vs_golden_fixture():
    """1,000 steps (the workload's length) against the oracle's 1,000-step
    run, committed as hashes, norms, masses and sampled values: within 1e-9
    relative, and bit for bit (equal SHA-256 of every field)."""
    with open(GOLDEN) as f:
        gold = json.load(f)
    m, (u, h, tr), (gh, gu, gt) = _l8_gpu(1000)
    assert _sha(h) == gold["initial"]["h"]["sha256"]  # the same seeded state
    ref = gold["checkpoints"]["1000"]
    for name, x in (("h", gh), ("u", gu), ("tracers", gt)):
        r = ref[name]
        idx = np.asarray(r["index"])
        val = np.asarray(r["value"])
        assert np.abs(x[idx] - val).max() <= 1e-9 * np.abs(val).max(), name
        assert abs(np.sqrt(np.sum(x * x)) - r["l2"]) <= 1e-9 * r["l2"], name
        assert abs(np.abs(x).max() - r["max"]) <= 1e-9 * r["max"], name
        assert _sha(x) == r["sha256"], f"{name}: not bit-identical to the oracle"
    area = m["areaCell"]
    mass = float(np.sum(gh * area))
    assert abs(mass / gold["initial"]["mass"] - 1) <= 1e-13
    assert abs(mass / ref["mass"] - 1) <= 1e-13


# ---- adjoint, level 6 --------------------------------------------------------

def test_adjoint_level6_200_steps():
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    level, steps, dt, nT, K, n_obs = 6, 200, 240.0, 1, 20, 1000
    m = mesh(level)
    C, E = m.nCells, m.nEdges
    u, h, tr = sw_workload_state(m, nT, seed=3)
    tr = tr.ravel()
    ctx = Context(m, nTracers=nT, tracer_scheme=TRACER_FCT, **FCT)
    ctx.sw_upload(h, u, tr, np.full(C, 4500.0))
    ctx.adj_init(K)
    x0 = ctx.state_arrays()
    for a, v in zip(x0, (u, h, tr)):
        a.upload(v)
    rng = np.random.default_rng(2024)
    obs = rng.choice(C, n_obs, replace=False)
    xp, xo = ctx.state_arrays(), ctx.state_arrays()
    for a, v, sc in zip(xp, (u, h, tr), (0.5, 2.0, 0.02)):
        a.upload(v + sc * rng.standard_normal(v.size))
    _say("adjoint: observations from a perturbed run")
    ctx.adj_forward(dt, steps, xp, xo)
    ctx.sync()
    h_obs, t_obs = xo[1].numpy()[obs], xo[2].numpy()[obs]

    def loss(x):
        hN, tN = x[1].numpy(), x[2].numpy()
        return float(np.sum((hN[obs] - h_obs) ** 2) + np.sum((tN[obs] - t_obs) ** 2))

    x1 = ctx.state_arrays()
    ctx.adj_forward(dt, steps, x0, x1)
    ctx.sync()
    hN, tN = x1[1].numpy(), x1[2].numpy()
    yb = ctx.state_arrays()
    sh, st = np.zeros(C), np.zeros(C)
    sh[obs] = 2 * (hN[obs] - h_obs)
    st[obs] = 2 * (tN[obs] - t_obs)
    for a, v in zip(yb, (np.zeros(E), sh, st)):
        a.upload(v)
    grad = ctx.state_arrays()
    _say("adjoint: sweep of 200 steps")
    ctx.adj_backward(yb, grad)
    ctx.sync()
    g = [a.numpy() for a in grad]
    assert all(np.isfinite(x).all() for x in g)

    # dot-product test over the whole run
    dx_h = [rng.standard_normal(E), rng.standard_normal(C), 0.01 * rng.standard_normal(C)]
    y_h = [rng.standard_normal(E), rng.standard_normal(C), rng.standard_normal(C)]
    dx, y, dxN, xb = (ctx.state_arrays() for _ in range(4))
    for a, v in zip(dx, dx_h):
        a.upload(v)
    for a, v in zip(y, y_h):
        a.upload(v)
    _say("adjoint: dot-product test")
    ctx.tlm_forward(dt, steps, x0, dx, dx1=dxN)
    ctx.adj_forward(dt, steps, x0)
    ctx.adj_backward(y, xb)
    ctx.sync()
    lhs = sum(float(np.dot(a.numpy(), v)) for a, v in zip(dxN, y_h))
    rhs = sum(float(np.dot(v, a.numpy())) for a, v in zip(xb, dx_h))
    _say(f"adjoint: <M dx, y> = {lhs:.16e}, <dx, M^T y> = {rhs:.16e}")
    assert abs(lhs - rhs) / abs(lhs) <= 1e-10

    # the tangent-linear model against central differences of the
    # frozen-limiter forward (the map it differentiates), eps = 1e-7, over
    # the segment the tape holds after a sweep (the first K steps of the
    # swept forward: here the unperturbed one), in a direction relative to
    # the state (u and the tracer scaled by themselves)
    eps = 1e-7
    d = [0.01 * u * dx_h[0], dx_h[1], tr * dx_h[2]]
    dd, dK, xK = ctx.state_arrays(), ctx.state_arrays(), ctx.state_arrays()
    for a, v in zip(dd, d):
        a.upload(v)
    ctx.tlm_forward(dt, K, x0, dd, x1=xK, dx1=dK)
    ctx.adj_forward(dt, steps, x0)
    ctx.adj_backward(y, xb)
    pK, mK = ctx.state_arrays(), ctx.state_arrays()
    for sgn, out in ((1, pK), (-1, mK)):
        xe = ctx.state_arrays()
        for a, v, dv in zip(xe, (u, h, tr), d):
            a.upload(v + sgn * eps * dv)
        ctx.adj_frozen_forward(K, xe, out)
    ctx.sync()
    for i, name in enumerate(("u", "h", "tracer")):
        tl = dK[i].numpy()
        fdv = (pK[i].numpy() - mK[i].numpy()) / (2 * eps)
        rel = np.linalg.norm(fdv - tl) / np.linalg.norm(tl)
        _say(f"adjoint: TLM vs frozen-limiter differences, {name}: {rel:.2e}")
        assert rel < 1e-4, name

    # the gradient against central differences of the loss, same eps and
    # direction, over all 200 steps of the plain forward, whose limiter is
    # not frozen: switches of the limiter inside [x - eps d, x + eps d] leave
    # a difference of about 1e-4 of the derivative
    jd = []
    for sgn in (1, -1):
        xe = ctx.state_arrays()
        for a, v, dv in zip(xe, (u, h, tr), d):
            a.upload(v + sgn * eps * dv)
        ctx.adj_forward(dt, steps, xe, x1)
        ctx.sync()
        jd.append(loss(x1))
    fd = (jd[0] - jd[1]) / (2 * eps)
    ga = sum(float(np.dot(gv, dv)) for gv, dv in zip(g, d))
    _say(f"adjoint: gradient . d = {ga:.10e}, central difference {fd:.10e}")
    assert abs(ga - fd) <= 1e-3 * abs(fd)
    ctx.close()


# ---- ensemble: 96 x L5 + 128 x L6 + 32 x L7 -----------------------------------

ENS_GROUPS = ((5, 96, 480.0), (6, 128, 240.0), (7, 32, 120.0))


def test_ensemble_256_members():
    from mpas_hip import BATCH_GRAPH, TRACER_FCT, Batch, Context, sw_workload_state
    from oracle.oracle import OracleMesh, OracleSW
    nT, steps = 2, 3
    ctxs = [Context(mesh(L), nTracers=nT, tracer_scheme=TRACER_FCT, **FCT)
            for L, _, _ in ENS_GROUPS]
    mesh_of, dts, states = [], [], []
    for g, (L, n, dt) in enumerate(ENS_GROUPS):
        for _ in range(n):
            i = len(mesh_of)
            rng = np.random.default_rng(100 + i)
            amp, u0 = rng.uniform(5.0, 20.0), rng.uniform(10.0, 30.0)
            u, h, tr = sw_workload_state(mesh(L), nT, seed=1 + i, u0=u0, bump_amp=amp)
            mesh_of.append(g)
            dts.append(dt)
            states.append((h, u, tr.ravel(), np.full(mesh(L).nCells, 4500.0)))
    b = Batch(ctxs, mesh_of, dts)
    assert b.n == 256 and b.groups() == 3
    for i, s in enumerate(states):
        b.upload(i, *s)
    _say("ensemble: autotuning (1 step per trial)")
    b.autotune(trial_steps=1)
    _say(f"ensemble: {b.get_config()}; {steps} graph steps")
    b.step(steps, BATCH_GRAPH)
    b.sync()
    # two members per group (the first and the last) against single runs
    first = [mesh_of.index(g) for g in range(3)]
    probe = sorted(set(first) | {len(mesh_of) - 1 - mesh_of[::-1].index(g) for g in range(3)})
    for i in probe:
        c = ctxs[mesh_of[i]]
        c.sw_upload(*states[i])
        c.rk4_step(dts[i], steps)
        for a, r in zip(b.download(i), c.sw_download()):
            assert bitwise_equal(a, r), i
    # the first level-5 member against the oracle
    i = first[0]
    h, u, tr, h_s = states[i]
    o = OracleSW(OracleMesh(mesh(5)), nT=nT, tracer_adv=2, **FCT)
    o.set("h", h)
    o.set("u", u)
    o.set("tracers", tr)
    o.set("h_s", h_s)
    o.init_diagnostics()
    o.run(steps, dts[i])
    for a, name in zip(b.download(i), ("h", "u", "tracers")):
        assert bitwise_equal(a, o.get(name)), name
    b.close()
    for c in ctxs:
        c.close()


# ---- batched adjoint ensemble: 64 x L7 + 64 x L5 ------------------------------

def test_batched_adjoint_ensemble_128_members():
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    nT, steps, K, n_obs = 2, 100, 10, 2000
    kw = dict(nTracers=nT, tracer_scheme=TRACER_FCT, **FCT)
    for L, n, dt in ((7, 64, 120.0), (5, 64, 480.0)):
        m = mesh(L)
        C, E = m.nCells, m.nEdges
        members = []
        for i in range(n):
            j = i + (64 if L == 5 else 0)
            rng = np.random.default_rng(700 + j)
            u, h, tr = sw_workload_state(m, nT, seed=300 + j, u0=rng.uniform(10.0, 30.0),
                                         bump_amp=rng.uniform(5.0, 20.0))
            obs = rng.choice(C, n_obs, replace=False)
            pert = (0.5 * rng.standard_normal(E), 2.0 * rng.standard_normal(C),
                    0.02 * rng.standard_normal(C * nT))
            members.append(((u, h, tr.ravel()), obs, pert))
        # a mesh's members as the levels of as few equal contexts as keep
        # every tape array (each stencil slot's part of the flux products
        # counting as one) below 4 GiB: one context of 64 per mesh
        cap = max(1, (2 ** 32 - 1) // (8 * E * nT))
        nctx = -(-n // cap)
        per_ctx = -(-n // nctx)
        _say(f"batched adjoint: level {L}, {n} members in {nctx} contexts of {per_ctx}")

        def pack(xs, w):
            return np.stack([x.reshape(-1, w) for x in xs], axis=1).ravel()

        for a0_ in range(0, n, per_ctx):
            part = members[a0_:a0_ + per_ctx]
            nM = len(part)
            X0 = [pack([p[0][q] for p in part], 1 if q < 2 else nT) for q in range(3)]
            XP = [pack([p[0][q] + p[2][q] for p in part], 1 if q < 2 else nT) for q in range(3)]
            ctx = Context(m, nLevels=nM, **kw)
            ctx.sw_upload(X0[1], X0[0], X0[2], np.full(C, 4500.0))
            ctx.adj_init(K)
            x0, xp, x1 = ctx.state_arrays(), ctx.state_arrays(), ctx.state_arrays()
            for a, v in zip(x0, X0):
                a.upload(v)
            for a, v in zip(xp, XP):
                a.upload(v)
            ctx.adj_forward(dt, steps, xp, x1)  # observations
            ctx.sync()
            ho, to = x1[1].numpy().reshape(C, nM), x1[2].numpy().reshape(C, nM, nT)
            ctx.adj_forward(dt, steps, x0, x1)
            ctx.sync()
            hN, tN = x1[1].numpy().reshape(C, nM), x1[2].numpy().reshape(C, nM, nT)
            sh, st = np.zeros((C, nM)), np.zeros((C, nM, nT))
            for q, (_, obs, _) in enumerate(part):
                sh[obs, q] = 2 * (hN[obs, q] - ho[obs, q])
                st[obs, q, :] = 2 * (tN[obs, q, :] - to[obs, q, :])
            yb, gb = ctx.state_arrays(), ctx.state_arrays()
            for a, v in zip(yb, (np.zeros(E * nM), sh.ravel(), st.ravel())):
                a.upload(v)
            ctx.adj_backward(yb, gb)
            ctx.sync()
            G = [gb[0].numpy().reshape(E, nM), gb[1].numpy().reshape(C, nM),
                 gb[2].numpy().reshape(C, nM, nT)]
            # sampled members alone: the first and the last of the context
            c1 = Context(m, **kw)
            c1.sw_upload(part[0][0][1], part[0][0][0], part[0][0][2], np.full(C, 4500.0))
            c1.adj_init(K)
            for q in (0, nM - 1):
                (u, h, tr), obs, _ = part[q]
                a0, a1, y, g = (c1.state_arrays() for _ in range(4))
                for a, v in zip(a0, (u, h, tr)):
                    a.upload(v)
                c1.adj_forward(dt, steps, a0, a1)
                c1.sync()
                h1, t1 = a1[1].numpy(), a1[2].numpy().reshape(C, nT)
                assert bitwise_equal(h1, hN[:, q].copy()), (L, q)
                s_h, s_t = np.zeros(C), np.zeros((C, nT))
                s_h[obs] = 2 * (h1[obs] - ho[obs, q])
                s_t[obs, :] = 2 * (t1[obs, :] - to[obs, q, :])
                for a, v in zip(y, (np.zeros(E), s_h, s_t.ravel())):
                    a.upload(v)
                c1.adj_backward(y, g)
                c1.sync()
                assert bitwise_equal(g[0].numpy(), G[0][:, q].copy()), (L, q, "u")
                assert bitwise_equal(g[1].numpy(), G[1][:, q].copy()), (L, q, "h")
                assert bitwise_equal(g[2].numpy(), G[2][:, q, :].ravel().copy()), (L, q, "tracers")
            c1.close()
            # the dot-product test on the multi-member tape
            rng = np.random.default_rng(5 + L)
            dx_h = [rng.standard_normal(E * nM), rng.standard_normal(C * nM),
                    0.01 * rng.standard_normal(C * nM * nT)]
            y_h = [rng.standard_normal(E * nM), rng.standard_normal(C * nM),
                   rng.standard_normal(C * nM * nT)]
            dx, y, dxN, xb = (ctx.state_arrays() for _ in range(4))
            for a, v in zip(dx, dx_h):
                a.upload(v)
            for a, v in zip(y, y_h):
                a.upload(v)
            ctx.tlm_forward(dt, steps, x0, dx, dx1=dxN)
            ctx.adj_forward(dt, steps, x0)
            ctx.adj_backward(y, xb)
            ctx.sync()
            lhs = sum(float(np.dot(a.numpy(), v)) for a, v in zip(dxN, y_h))
            rhs = sum(float(np.dot(v, a.numpy())) for a, v in zip(xb, dx_h))
            _say(f"batched adjoint: level {L} dot-product rel {abs(lhs - rhs) / abs(lhs):.2e}")
            assert abs(lhs - rhs) / abs(lhs) <= 1e-10
            ctx.close()
