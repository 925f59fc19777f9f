#!/usr/bin/env python3
"""RK4 cell-steps/s of the single-instance drop-in workload on one MI355X.

Workload: level-8 icosahedral SCVT mesh (655,362 cells, 1,966,080 edges,
1,310,720 vertices), FP64 shallow-water RK4 with dt = 60 s, balanced
zonal flow plus a seeded 10 m Gaussian bump, 4 passive tracers (seeded
sums of 8 Gaussian blobs) advected with FCT, del2 tracer diffusion 1e3 m^2/s
and del2 viscosity 1e5 m^2/s.  One step = one full RK4 step (4 stages) over
the whole mesh with the state resident in HBM.

Prints one JSON line.  `roofline` is for the kernel with the largest share
of the step (HIP events around each launch, compulsory bytes per launch);
`cpu_baseline` times the oracle (the C++ restatement of the reference's
task loops, one core) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mpas-o-flecsi-2.0_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "RK4 cell-steps/sec + %HBM peak; adjoint/forward cost ratio at ≤1e-10 rel err"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def model_bytes_per_step(ctx, mesh, n_tr):
    """Counted compulsory bytes of one RK4 step: static mesh arrays once per
    stage plus 18 touches of the prognostic state (u, h, tracers)."""
    C, E, V = mesh.nCells, mesh.nEdges, mesh.nVertices
    s_eoe = int(mesh["nEdgesOnEdge"].sum())
    s_eoc = int(mesh["nEdgesOnCell"].sum())
    s_adv = int(ctx.setup_field("nAdvCellsForEdge").sum())
    stage = (36 * E + 12 * s_eoe + 20 * C + 4 * s_eoc + 64 * V + 8 * E
             + 20 * s_adv + 4 * E + 4 * s_eoc)
    state = 18 * 8 * (C * (1 + n_tr) + E)
    return 4 * stage + state


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh, u, h, tr, h_s, dt, n_tr, steps_1=2, steps_all=4, runs=5):
    """The oracle (the C++ restatement of the reference's task loops, g++ -O2)
    on the host cores, on a bounded sample of the workload: `runs` runs of
    `steps_1` steps on one thread (the reference's serial single-rank loops)
    and of `steps_all` steps on every core the box gives this process (its
    gather loops threaded, its scatter loops serial: the same bits), each
    from the same state after one untimed setup; medians of the runs."""
    from oracle.oracle import OracleMesh, OracleSW, set_threads
    t0 = time.perf_counter()
    om = OracleMesh(mesh, nL=1)
    setup_s = time.perf_counter() - t0
    cores = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)))

    def rate(threads, steps):
        set_threads(threads)
        o = OracleSW(om, nL=1, nT=n_tr, tracer_adv=2, visc2=1e5, kappa2=1e3)
        o.set("h", h)
        o.set("u", u)
        o.set("tracers", tr)
        o.set("h_s", h_s)
        o.init_diagnostics()
        return mesh.nCells * steps / o.run(steps, dt)

    one = [rate(1, steps_1) for _ in range(runs)]
    many = [rate(cores, steps_all) for _ in range(runs)]
    set_threads(1)
    return {"value": float(np.median(many)), "unit": "cell-steps/s", "cores": cores,
            "kind": "port",
            "sample": f"median of {runs} runs of {steps_all} RK4 steps of the same level-8 "
                      f"workload on the oracle (C++ restatement of the reference task loops, "
                      f"g++ -O2), gather loops on {cores} OpenMP threads, scatter loops serial; "
                      f"step loop only, mesh setup {setup_s:.1f} s untimed",
            "one_core": {"value": float(np.median(one)), "cores": 1,
                         "sample": f"median of {runs} runs of {steps_1} steps, one thread"},
            "runs_all_cores": many, "runs_one_core": one,
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def adjoint_bench(level=6, steps=200, dt=240.0, n_tr=1, k=20, n_obs=1000, reps=3):
    """Gradient of an observation loss with respect to the initial h, u and
    tracer (level-6 mesh, 200 steps of dt = 240 s, 1 FCT tracer,
    checkpoints every 20 steps).  Loss = sum over 1,000 seeded cells of the
    squared differences of h and the tracer at the final step from
    observations taken from a seeded-perturbed forward run.  Cost ratios:
    adjoint sweep / plain forward, and (forward with checkpoints + adjoint
    sweep) / plain forward; accuracy by the dot-product test
    <M dx, y> = <dx, M^T y> over the full 200 steps."""
    from mpas_hip import TRACER_FCT, Context, Mesh, sw_workload_state
    mesh = Mesh(level)
    u, h, tr = sw_workload_state(mesh, n_tr, seed=3)
    tr = tr.ravel()
    ctx = Context(mesh, nTracers=n_tr, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    ctx.sw_upload(h, u, tr, np.full(mesh.nCells, 4500.0))
    ctx.adj_init(k)
    # tile and schedule for this mesh, on half the run (5 segments: the
    # replays beside the sweeps settle as in the full run)
    adj_trials = ctx.adj_autotune(dt, 5 * k)
    adj_cfg = ctx.adj_get_config()
    x0 = ctx.state_arrays()
    for a, v in zip(x0, (u, h, tr)):
        a.upload(v)
    rng = np.random.default_rng(2024)
    obs = rng.choice(mesh.nCells, n_obs, replace=False)
    # observations from a forward run of a perturbed initial state
    xp = ctx.state_arrays()
    for a, v, sc in zip(xp, (u, h, tr), (0.5, 2.0, 0.02)):
        a.upload(v + sc * rng.standard_normal(v.size))
    xo = ctx.state_arrays()
    ctx.adj_forward(dt, steps, xp, xo)
    ctx.sync()
    h_obs, tr_obs = xo[1].numpy()[obs], xo[2].numpy()[obs]

    def timed(fn):
        best = float("inf")
        for _ in range(reps):
            ctx.sync()
            t0 = time.perf_counter()
            fn()
            ctx.sync()
            best = min(best, time.perf_counter() - t0)
        return best

    # plain forward: the fused step from the same state
    def plain():
        ctx.sw_upload(h, u, tr, np.full(mesh.nCells, 4500.0))
        ctx.sync()
    t_fwd = float("inf")
    for _ in range(reps):
        plain()
        t0 = time.perf_counter()
        ctx.rk4_step(dt, steps)
        ctx.sync()
        t_fwd = min(t_fwd, time.perf_counter() - t0)
    x1 = ctx.state_arrays()
    t_adj_fwd = timed(lambda: ctx.adj_forward(dt, steps, x0, x1))
    hN, trN = x1[1].numpy(), x1[2].numpy()
    J = float(np.sum((hN[obs] - h_obs) ** 2) + np.sum((trN[obs] - tr_obs) ** 2))
    seed_u = np.zeros(mesh.nEdges)
    seed_h = np.zeros(mesh.nCells)
    seed_t = np.zeros(mesh.nCells * n_tr)
    seed_h[obs] = 2 * (hN[obs] - h_obs)
    seed_t[obs] = 2 * (trN[obs] - tr_obs)
    yb = ctx.state_arrays()
    for a, v in zip(yb, (seed_u, seed_h, seed_t)):
        a.upload(v)
    g = ctx.state_arrays()
    t_bwd = timed(lambda: ctx.adj_backward(yb, g))
    grad = [a.numpy() for a in g]
    # dot-product test over the full run: random dx, y
    dx_h = [rng.standard_normal(mesh.nEdges), rng.standard_normal(mesh.nCells),
            0.01 * rng.standard_normal(mesh.nCells * n_tr)]
    y_h = [rng.standard_normal(mesh.nEdges), rng.standard_normal(mesh.nCells),
           rng.standard_normal(mesh.nCells * n_tr)]
    dx, y, dxN, xb = (ctx.state_arrays() for _ in range(4))
    for a, v in zip(dx, dx_h):
        a.upload(v)
    for a, v in zip(y, y_h):
        a.upload(v)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.tlm_forward(dt, steps, x0, dx, dx1=dxN)
    ctx.sync()
    t_tlm = time.perf_counter() - t0
    ctx.adj_forward(dt, steps, x0)
    ctx.adj_backward(y, xb)
    ctx.sync()
    lhs = sum(float(np.dot(a.numpy(), v)) for a, v in zip(dxN, y_h))
    rhs = sum(float(np.dot(v, a.numpy())) for a, v in zip(xb, dx_h))
    dot_rel = abs(lhs - rhs) / abs(lhs)
    # central differences of the loss (the plain forward, limiter not frozen)
    eps = 1e-7
    d_h = [0.01 * u * dx_h[0], dx_h[1], tr * dx_h[2]]
    jd = []
    for sgn in (1, -1):
        xe = ctx.state_arrays()
        for a, v, d in zip(xe, (u, h, tr), d_h):
            a.upload(v + sgn * eps * d)
        ctx.adj_forward(dt, steps, xe, x1)
        ctx.sync()
        he, te = x1[1].numpy(), x1[2].numpy()
        jd.append(float(np.sum((he[obs] - h_obs) ** 2) + np.sum((te[obs] - tr_obs) ** 2)))
    fd = (jd[0] - jd[1]) / (2 * eps)
    ga = sum(float(np.dot(gv, d)) for gv, d in zip(grad, d_h))
    ctx.close()
    return {
        "workload": f"icosahedral SCVT level {level} ({mesh.nCells} cells), FP64 SW RK4 + FCT, "
                    f"{steps} steps dt={dt:g} s, {n_tr} tracer, checkpoint every {k} steps, "
                    f"loss over {n_obs} seeded cells (h and tracer) vs perturbed-run observations",
        "tuned_config": adj_cfg, "autotune_trials": adj_trials,
        "forward_ms": 1e3 * t_fwd, "forward_with_checkpoints_ms": 1e3 * t_adj_fwd,
        "backward_ms": 1e3 * t_bwd, "tlm_ms": 1e3 * t_tlm,
        # the adjoint sweep (segment replays from the checkpoints included)
        # against the forward pass it differentiates, and the whole gradient
        # (forward with checkpoints + adjoint sweep) against the forward
        "adjoint_forward_cost_ratio": t_bwd / t_fwd,
        "gradient_total_cost_ratio": (t_adj_fwd + t_bwd) / t_fwd,
        "loss": J, "dot_product_rel_err": dot_rel,
        "dot_product": {"tlm_side": lhs, "adjoint_side": rhs},
        "fd_check": {"eps": eps, "adjoint": ga, "central_difference": fd,
                     "rel_diff": abs(ga - fd) / abs(fd),
                     "note": "plain forward: the limiter is not frozen in the differences"},
    }


def ensemble_bench(steps=500, seq_steps=20, n_tr=2, tune_steps=10,
                   groups=((5, 96, 480.0), (6, 128, 240.0), (7, 32, 120.0)), trace_steps=0):
    """256-member ensemble on meshes of uneven size (96 x level 5 at dt 480 s,
    128 x level 6 at dt 240 s, 32 x level 7 at dt 120 s), 2 FCT tracers, each
    member with a seeded bump amplitude in [5, 20] m and u0 in [10, 30] m/s.
    Batched: one launch per phase for all members (members of a mesh stored
    member-minor), the step replayed from a hipGraph, the launch
    configuration from the autotuner.  Baseline: each member run alone on its
    mesh's context (mpas_rk4_step: the single-run launches), member after
    member, timed over seq_steps steps per member (its cost per step is
    constant).  The batched members equal the single runs bit for bit
    (checked on one member per mesh).

    trace_steps > 0: a kernel-trace run instead, with the default launch
    configuration (no autotuning) and trace_steps direct-launch steps (one
    launch per phase and layout, no graph), nothing else."""
    from mpas_hip import (BATCH_GRAPH, BATCH_LAUNCH, BATCH_PERSISTENT, BATCH_TUNED, TRACER_FCT,
                          Batch, Context, Mesh, sw_workload_state)
    t0 = time.perf_counter()
    meshes = [Mesh(L) for L, _, _ in groups]
    ctxs = [Context(m, nTracers=n_tr, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
            for m in meshes]
    mesh_of, dts, states = [], [], []
    for g, (L, n, dt) in enumerate(groups):
        for _ in range(n):
            i = len(mesh_of)
            rng = np.random.default_rng(100 + i)
            amp, u0 = rng.uniform(5.0, 20.0), rng.uniform(10.0, 30.0)
            u, h, tr = sw_workload_state(meshes[g], n_tr, seed=1 + i, u0=u0, bump_amp=amp)
            mesh_of.append(g)
            dts.append(dt)
            states.append((h, u, tr.ravel(), np.full(meshes[g].nCells, 4500.0)))
    b = Batch(ctxs, mesh_of, dts)

    def load():
        for i, st in enumerate(states):
            b.upload(i, *st)

    load()
    setup_s = time.perf_counter() - t0
    log(f"  ensemble: {len(mesh_of)} members in {b.groups()} groups set up in {setup_s:.1f} s")
    if trace_steps > 0:
        b.step(2, BATCH_LAUNCH)
        b.sync()
        t = time.perf_counter()
        b.step(trace_steps, BATCH_LAUNCH)
        b.sync()
        ms = 1e3 * (time.perf_counter() - t) / trace_steps
        cfg = b.get_config()
        b.close()
        for c in ctxs:
            c.close()
        return {"trace_steps": trace_steps, "ms_per_step_batched_launch": ms, "config": cfg}
    trials = b.autotune(trial_steps=tune_steps)
    cfg = b.get_config()
    cfg["mesh_order"] = ["given (Hilbert)", "reverse Cuthill-McKee", "Morton"][b.get_order()]
    order_trials = {["given (Hilbert)", "reverse Cuthill-McKee", "Morton"][t["mesh_order"]]:
                    t["ms_per_step"] for t in trials[:3] if "mesh_order" in t}
    log(f"  ensemble: autotuned {cfg} over {len(trials)} trials; mesh orders {order_trials}")

    def run(nsteps, mode):
        load()
        b.sync()
        t = time.perf_counter()
        b.step(nsteps, mode)
        b.sync()
        return time.perf_counter() - t

    run(2, BATCH_GRAPH)  # warm (graph capture)
    t_graph = run(steps, BATCH_GRAPH)
    log(f"  ensemble: {steps} batched steps (graph) {1e3 * t_graph / steps:.3f} ms/step")
    t_launch = run(steps, BATCH_LAUNCH)
    run(2, BATCH_PERSISTENT)
    t_pers = run(steps, BATCH_PERSISTENT)
    log(f"  ensemble: {steps} batched steps (persistent launch) "
        f"{1e3 * t_pers / steps:.3f} ms/step")
    tuned = b.tuned_mode()
    t_tuned = run(steps, BATCH_TUNED)
    mode_name = {BATCH_GRAPH: "graph", BATCH_PERSISTENT: "persistent launch"}[tuned]
    log(f"  ensemble: tuned mode {mode_name}: {1e3 * t_tuned / steps:.3f} ms/step")
    # baseline: every member alone, seq_steps steps each; bitwise check on
    # one member per mesh against the batched run of the same length
    probe = {mesh_of.index(g) for g in range(len(groups))}
    run(seq_steps, BATCH_TUNED)
    got = {i: b.download(i) for i in probe}
    t_seq, same = 0.0, True
    for i, st in enumerate(states):
        c = ctxs[mesh_of[i]]
        c.sw_upload(*st)
        c.sync()
        t = time.perf_counter()
        c.rk4_step(dts[i], seq_steps)
        c.sync()
        t_seq += time.perf_counter() - t
        if i in probe:
            same = same and all(np.array_equal(x.view(np.int64), y.view(np.int64))
                                for x, y in zip(got[i], c.sw_download()))
    log(f"  ensemble: members alone {1e3 * t_seq / seq_steps:.3f} ms/step")
    cells = sum(meshes[g].nCells for g in mesh_of)
    # counted bytes: static mesh arrays once per mesh per stage, state per member
    model = 0.0
    for g, (L, n, _) in enumerate(groups):
        per = model_bytes_per_step(ctxs[g], meshes[g], n_tr)
        state = 18 * 8 * (meshes[g].nCells * (1 + n_tr) + meshes[g].nEdges)
        model += (per - state) + n * state
    b.close()
    for c in ctxs:
        c.close()
    ms_graph = 1e3 * t_graph / steps
    ms_tuned = 1e3 * t_tuned / steps
    ms_seq = 1e3 * t_seq / seq_steps
    return {
        "workload": "ensemble of 256 members: 96 x level 5 (10,242 cells, dt 480 s), 128 x "
                    "level 6 (40,962 cells, dt 240 s), 32 x level 7 (163,842 cells, dt 120 s); "
                    f"FP64 SW RK4 + FCT, {n_tr} tracers, seeded bump 5-20 m and u0 10-30 m/s",
        "members": len(mesh_of), "member_cells": cells, "steps": steps,
        "ms_per_step_batched_graph": ms_graph,
        "ms_per_step_batched_launch": 1e3 * t_launch / steps,
        "ms_per_step_persistent_launch": 1e3 * t_pers / steps,
        "tuned_mode": mode_name, "ms_per_step_tuned": ms_tuned,
        "ms_per_step_sequential": ms_seq, "sequential_steps_timed": seq_steps,
        "speedup_vs_sequential": ms_seq / ms_tuned,
        "member_cell_steps_per_s": cells * steps / t_tuned,
        "hbm_frac_model": {"bytes_per_step": model,
                           "achieved_GBs": model / (ms_tuned * 1e-3) / 1e9,
                           "frac": model / (ms_tuned * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "bitwise_equal_to_single_runs": bool(same),
        "autotuned_config": cfg, "autotune_trials": len(trials),
        "mesh_order_trials_ms_per_step": order_trials,
        "setup_s": setup_s,
    }


def adjoint_ensemble_bench(steps=100, n_tr=2, k=10, n_obs=2000,
                           groups=((7, 64, 120.0), (5, 64, 480.0)), sample=2, baseline=True):
    """Batched adjoint ensemble: 64 members on the level-7 mesh and 64 on the
    level-5 mesh, 2 FCT tracers, 100 steps, checkpoints every 10 steps, each
    member with its own seeded bump and its own seeded set of 2,000
    observation cells (observations from a seeded-perturbed run of the
    member).  The members of a mesh are the levels of one context, one level
    per lane; a mesh's members are split over as few equal contexts as keep
    every tape array (each stencil slot's part of the per-slot flux products
    counting as one) below the 4 GiB the kernels address (the largest holds
    edges x levels x tracers); forward with checkpoints and
    adjoint sweep of every context run back to back.  Baseline: every member
    alone (its own single-level adjoint, the context tuned likewise), all of
    them timed; the first `sample` members per mesh also compare their
    gradients with the batch's, bit for bit.  baseline=False skips the
    baseline and the autotuner (for a kernel trace of the batched part:
    workgroups of 512, replays beside the sweep, no graphs)."""
    from mpas_hip import TRACER_FCT, Context, Mesh, sw_workload_state
    kw = dict(nTracers=n_tr, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    batches, singles, total = [], [], 0
    t_setup = time.perf_counter()
    for L, n, dt in groups:
        mesh = Mesh(L)
        C, E = mesh.nCells, mesh.nEdges
        members = []
        for i in range(n):
            j = total + i
            rng = np.random.default_rng(700 + j)
            u, h, tr = sw_workload_state(mesh, n_tr, seed=300 + j, u0=rng.uniform(10.0, 30.0),
                                         bump_amp=rng.uniform(5.0, 20.0))
            obs = rng.choice(C, n_obs, replace=False)
            pert = (0.5 * rng.standard_normal(E), 2.0 * rng.standard_normal(C),
                    0.02 * rng.standard_normal(C * n_tr))
            members.append(((u, h, tr.ravel()), obs, pert))
        total += n
        cap = max(1, (2 ** 32 - 1) // (8 * E * n_tr))
        nctx = -(-n // cap)
        per_ctx = -(-n // nctx)
        for a in range(0, n, per_ctx):
            part = members[a:a + per_ctx]
            nM = len(part)
            pack = lambda xs, w: np.stack([x.reshape(-1, w) for x in xs], axis=1).ravel()
            X0 = [pack([m_[0][0] for m_ in part], 1), pack([m_[0][1] for m_ in part], 1),
                  pack([m_[0][2] for m_ in part], n_tr)]
            XP = [pack([m_[0][q] + m_[2][q] for m_ in part], 1 if q < 2 else n_tr)
                  for q in range(3)]
            ctx = Context(mesh, nLevels=nM, **kw)
            ctx.sw_upload(X0[1], X0[0], X0[2], np.full(C, 4500.0))
            ctx.adj_init(k)
            if baseline:
                ctx.adj_autotune(dt, 2 * k)  # tile and schedule for this mesh and member count
            else:  # kernel-trace runs: the profiler's kernel tracing fails on graph capture
                ctx.adj_config(512, 512, True, False)
            x0, xp, x1 = ctx.state_arrays(), ctx.state_arrays(), ctx.state_arrays()
            for d, v in zip(x0, X0):
                d.upload(v)
            for d, v in zip(xp, XP):
                d.upload(v)
            ctx.adj_forward(dt, steps, xp, x1)  # observations
            ctx.sync()
            ho = x1[1].numpy().reshape(C, nM)
            to = x1[2].numpy().reshape(C, nM, n_tr)
            batches.append(dict(ctx=ctx, mesh=mesh, L=L, dt=dt, part=part, nM=nM, x0=x0, x1=x1,
                                ho=ho, to=to, yb=ctx.state_arrays(), g=ctx.state_arrays()))
        singles.append((mesh, dt, members, n))
    t_setup = time.perf_counter() - t_setup

    def seeds(b, hN, tN):
        C, nM = b["mesh"].nCells, b["nM"]
        sh, st_ = np.zeros((C, nM)), np.zeros((C, nM, n_tr))
        loss = 0.0
        for q, (_, obs, _) in enumerate(b["part"]):
            dh = hN[obs, q] - b["ho"][obs, q]
            dT = tN[obs, q, :] - b["to"][obs, q, :]
            sh[obs, q] = 2 * dh
            st_[obs, q, :] = 2 * dT
            loss += float(np.sum(dh ** 2) + np.sum(dT ** 2))
        return sh.ravel(), st_.ravel(), loss

    # batched: forward with checkpoints, then the adjoint sweep, every context
    for b in batches:
        b["ctx"].sync()
    t0 = time.perf_counter()
    for b in batches:
        b["ctx"].adj_forward(b["dt"], steps, b["x0"], b["x1"])
    for b in batches:
        b["ctx"].sync()
    t_fwd = time.perf_counter() - t0
    loss = 0.0
    for b in batches:  # the loss seeds (host work, not timed)
        C, nM = b["mesh"].nCells, b["nM"]
        sh, st_, l_ = seeds(b, b["x1"][1].numpy().reshape(C, nM),
                            b["x1"][2].numpy().reshape(C, nM, n_tr))
        loss += l_
        b["yb"][0].upload(np.zeros(b["mesh"].nEdges * nM))
        b["yb"][1].upload(sh)
        b["yb"][2].upload(st_)
        b["ctx"].sync()
    t0 = time.perf_counter()
    for b in batches:
        b["ctx"].adj_backward(b["yb"], b["g"])
    for b in batches:
        b["ctx"].sync()
    t_bwd = time.perf_counter() - t0
    # baseline: every member alone (its own single-level context, tuned
    # likewise), and the first `sample` members' gradients against the batch
    t_seq, same, tuned_single = 0.0, True, {}
    for (mesh, dt, all_members, n), b in zip(singles if baseline else [],
                                             [next(x for x in batches if x["L"] == g[0])
                                              for g in groups]):
        C, E = mesh.nCells, mesh.nEdges
        c1 = Context(mesh, **kw)
        (u, h, tr), _, _ = all_members[0]
        c1.sw_upload(h, u, tr, np.full(C, 4500.0))
        c1.adj_init(k)
        c1.adj_autotune(dt, 2 * k)
        tuned_single[f"level {mesh.level}"] = c1.adj_get_config()
        per = []
        gb = [a.numpy() for a in b["g"]]
        for q, ((u, h, tr), obs, _) in enumerate(all_members):
            if q >= sample:  # timing only: the members of the other contexts
                a0, a1, y, g = (c1.state_arrays() for _ in range(4))
                for d_, v in zip(a0, (u, h, tr)):
                    d_.upload(v)
                c1.sync()
                t0 = time.perf_counter()
                c1.adj_forward(dt, steps, a0, a1)
                c1.sync()
                tf = time.perf_counter() - t0
                for d_, v in zip(y, (np.zeros(E), np.ones(C), np.ones(C * n_tr))):
                    d_.upload(v)
                c1.sync()
                t1 = time.perf_counter()
                c1.adj_backward(y, g)
                c1.sync()
                per.append(tf + time.perf_counter() - t1)
                for arr in (a0, a1, y, g):
                    for d_ in arr:
                        d_.free()
                continue
            a0, a1, y, g = (c1.state_arrays() for _ in range(4))
            for d, v in zip(a0, (u, h, tr)):
                d.upload(v)
            c1.sync()
            t0 = time.perf_counter()
            c1.adj_forward(dt, steps, a0, a1)
            c1.sync()
            tf = time.perf_counter() - t0
            hN, tN = a1[1].numpy(), a1[2].numpy().reshape(C, n_tr)
            sh, st_ = np.zeros(C), np.zeros((C, n_tr))
            sh[obs] = 2 * (hN[obs] - b["ho"][obs, q])
            st_[obs, :] = 2 * (tN[obs, :] - b["to"][obs, q, :])
            for d, v in zip(y, (np.zeros(E), sh, st_.ravel())):
                d.upload(v)
            c1.sync()
            t1 = time.perf_counter()
            c1.adj_backward(y, g)
            c1.sync()
            per.append(tf + time.perf_counter() - t1)
            nM = b["nM"]
            same &= np.array_equal(g[0].numpy().view(np.uint64),
                                   gb[0].reshape(E, nM)[:, q].copy().view(np.uint64))
            same &= np.array_equal(g[1].numpy().view(np.uint64),
                                   gb[1].reshape(C, nM)[:, q].copy().view(np.uint64))
            same &= np.array_equal(g[2].numpy().view(np.uint64),
                                   gb[2].reshape(C, nM, n_tr)[:, q, :].ravel().copy()
                                   .view(np.uint64))
        c1.close()
        t_seq += float(np.sum(per))
    tuned_ctx = [dict(level=b["L"], members=b["nM"], **b["ctx"].adj_get_config())
                 for b in batches]
    for b in batches:
        b["ctx"].close()
    t_batch = t_fwd + t_bwd
    return {
        "workload": "batched adjoint ensemble: " + ", ".join(
            f"{n} members on level {L} (dt {dt:g} s)" for L, n, dt in groups) +
            f"; FP64 SW RK4 + FCT, {n_tr} tracers, {steps} steps, checkpoint every {k} steps, "
            f"{n_obs} seeded observation cells per member; members of a mesh are the levels "
            f"of one context (split in equal contexts where a tape array would pass 4 GiB)",
        "members": total, "contexts": len(batches), "setup_s": t_setup,
        "tuned_config_per_context": tuned_ctx,
        "tuned_config_single_runs": tuned_single,
        "forward_with_checkpoints_ms": 1e3 * t_fwd, "backward_ms": 1e3 * t_bwd,
        "batched_ms": 1e3 * t_batch, "sequential_ms": 1e3 * t_seq if baseline else None,
        "sequential_members_timed": total if baseline else 0,
        "speedup_vs_sequential": t_seq / t_batch if baseline else None, "loss": loss,
        "sampled_gradients_bitwise_equal_to_single_runs": bool(same) if baseline else None,
    }


def copy_calibration(nbytes):
    """k_copy over nbytes, 10 launches: reads and writes nbytes each."""
    from mpas_hip import Context, Mesh
    nbytes -= nbytes % 16
    ctx = Context(Mesh(0))
    a, b = ctx.alloc(nbytes // 8), ctx.alloc(nbytes // 8)
    ctx.copy_probe(b, a, nbytes)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(10):
        ctx.copy_probe(b, a, nbytes)
    ctx.sync()
    dt = (time.perf_counter() - t0) / 10
    print(json.dumps({"copy_bytes": nbytes, "launches": 11, "s_per_launch": dt,
                      "GBs_read_plus_write": 2 * nbytes / dt / 1e9}), flush=True)
    ctx.close()


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--tracers", type=int, default=4)
    ap.add_argument("--dt", type=float, default=60.0)
    ap.add_argument("--prof-steps", type=int, default=50,
                    help="steps of the per-kernel event-timed pass")
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="oracle steps per one-thread run of the CPU baseline (about 1.5 s "
                         "each); the all-cores runs take twice as many")
    ap.add_argument("--repeats", type=int, default=5,
                    help="further timed blocks of --steps steps for the median")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint measurement (level 6, 200 steps, 1 FCT tracer)")
    ap.add_argument("--adjoint-only", action="store_true")
    ap.add_argument("--adj-steps", type=int, default=200,
                    help="steps of the adjoint measurement (its workload has 200)")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the 256-member ensemble measurement")
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--ens-steps", type=int, default=500,
                    help="batched steps of the ensemble measurement (its workload has 500)")
    ap.add_argument("--ens-seq-steps", type=int, default=20,
                    help="steps per member of the member-after-member baseline")
    ap.add_argument("--ens-trace-steps", type=int, default=0,
                    help="with --ensemble-only: only this many direct-launch batched steps "
                         "in the default configuration (for a kernel trace)")
    ap.add_argument("--no-adjoint-ensemble", action="store_true",
                    help="skip the batched adjoint ensemble (64 x level 7 + 64 x level 5)")
    ap.add_argument("--adjoint-ensemble-only", action="store_true")
    ap.add_argument("--adj-ens-no-baseline", action="store_true",
                    help="with --adjoint-ensemble-only: skip the member-alone baseline "
                         "(for a kernel trace of the batched part)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel k_copy over BYTES, 10 times "
                         "(the known-size reference for HBM counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU step "
              "(multi-GPU sharding is a design note, not an implemented path)", file=sys.stderr)
        sys.exit(2)

    from mpas_hip import TRACER_FCT, Context, Mesh, sw_workload_state

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_bench(steps=args.adj_steps)), flush=True)
        return
    if args.adjoint_ensemble_only:
        print(json.dumps(adjoint_ensemble_bench(baseline=not args.adj_ens_no_baseline)),
              flush=True)
        return
    if args.ensemble_only:
        print(json.dumps(ensemble_bench(steps=args.ens_steps, seq_steps=args.ens_seq_steps,
                                        trace_steps=args.ens_trace_steps)), flush=True)
        return

    t0 = time.perf_counter()
    mesh = Mesh(args.level)
    u, h, tr = sw_workload_state(mesh, args.tracers, seed=1)
    h_s = np.full(mesh.nCells, 4500.0)
    log(f"mesh L{args.level}: {mesh.nCells} cells, {mesh.nEdges} edges "
        f"({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    ctx = Context(mesh, nTracers=args.tracers, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    ctx.sw_upload(h, u, tr, h_s)
    log(f"context + upload {time.perf_counter() - t0:.1f} s")

    ctx.rk4_step(args.dt, args.warmup)
    ctx.sync()
    log(f"warmup {args.warmup} steps done")
    t0 = time.perf_counter()
    ctx.rk4_step(args.dt, args.steps)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    value = mesh.nCells * args.steps / elapsed
    log(f"timed {args.steps} steps: {ms_per_step:.4f} ms/step")
    blocks = []
    for _ in range(args.repeats):
        t0 = time.perf_counter()
        ctx.rk4_step(args.dt, args.steps)
        ctx.sync()
        blocks.append(1e3 * (time.perf_counter() - t0) / args.steps)
    if blocks:
        log(f"{args.repeats} more blocks of {args.steps} steps: median "
            f"{float(np.median(blocks)):.4f} ms/step")

    # per-kernel HIP-event pass on the context's stream
    ctx.profile(True)
    t0 = time.perf_counter()
    ctx.rk4_step(args.dt, args.prof_steps)
    ctx.sync()
    prof_ms_step = 1e3 * (time.perf_counter() - t0) / max(1, args.prof_steps)
    rep = ctx.profile_report()
    ctx.profile(False)
    tot = sum(r["total_ms"] for r in rep)
    for r in sorted(rep, key=lambda r: -r["total_ms"]):
        avg = r["total_ms"] / r["launches"]
        log(f"  {r['kernel']:16s} {r['launches']:6d} launches  avg {1e3 * avg:8.1f} us  "
            f"{100 * r['total_ms'] / tot:5.1f}%  {r['bytes_per_launch'] / avg / 1e6:7.0f} GB/s")
    dom = max(rep, key=lambda r: r["total_ms"])
    avg_s = dom["total_ms"] / dom["launches"] / 1e3
    achieved = dom["bytes_per_launch"] / avg_s / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get(dom["kernel"], {}).get("hbm_bytes_per_launch")

    h1, _, _ = ctx.sw_download()
    area = mesh["areaCell"]
    mass_drift = float(np.sum(h1 * area) / np.sum(h * area) - 1.0)
    model = model_bytes_per_step(ctx, mesh, args.tracers)
    ctx.close()

    out = {
        "metric": METRIC, "value": value, "unit": "cell-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"icosahedral SCVT level {args.level} ({mesh.nCells} cells, "
                               f"{mesh.nEdges} edges, {mesh.nVertices} vertices), FP64 SW RK4 "
                               f"dt={args.dt:g} s, {args.tracers} FCT tracers, kappa2=1e3, "
                               f"visc2=1e5",
                   "cells": mesh.nCells, "edges": mesh.nEdges, "vertices": mesh.nVertices,
                   "tracers": args.tracers, "dt": args.dt, "parallelism": "single GPU"},
        "hbm_frac_model": {"bytes_per_cell_step": model / mesh.nCells,
                           "achieved_GBs": model * args.steps / elapsed / 1e9,
                           "frac": model * args.steps / elapsed / 1e9 / HBM_PEAK_GBS},
        "roofline": {"bound": "hbm", "kernel": dom["kernel"], "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "bytes_per_launch": dom["bytes_per_launch"],
                     "avg_launch_us": avg_s * 1e6, "profiled_ms_per_step": prof_ms_step},
        "mass_drift": mass_drift,
        "repeats": {"blocks": len(blocks), "steps_each": args.steps,
                    "ms_per_step_median": float(np.median(blocks)) if blocks else None,
                    "ms_per_step_min": min(blocks) if blocks else None,
                    "ms_per_step_max": max(blocks) if blocks else None},
    }
    if not args.no_adjoint:
        log("adjoint (level 6, 200 steps) ...")
        out["adjoint"] = adjoint_bench(steps=args.adj_steps)
        log(f"  adjoint/forward {out['adjoint']['adjoint_forward_cost_ratio']:.3f}, "
            f"gradient/forward {out['adjoint']['gradient_total_cost_ratio']:.3f}, "
            f"dot-product rel err {out['adjoint']['dot_product_rel_err']:.2e}")
    if not args.no_ensemble:
        log("ensemble (256 members, 500 steps) ...")
        out["ensemble"] = ensemble_bench(steps=args.ens_steps, seq_steps=args.ens_seq_steps)
        log(f"  {out['ensemble']['ms_per_step_tuned']:.3f} ms/step batched, "
            f"{out['ensemble']['ms_per_step_sequential']:.3f} sequential, speedup "
            f"{out['ensemble']['speedup_vs_sequential']:.1f}")
    if not args.no_adjoint_ensemble:
        log("batched adjoint ensemble (64 x L7 + 64 x L5, 100 steps) ...")
        out["adjoint_ensemble"] = adjoint_ensemble_bench()
        ae = out["adjoint_ensemble"]
        log(f"  {ae['batched_ms']:.0f} ms batched, {ae['sequential_ms']:.0f} ms "
            f"sequential, speedup {ae['speedup_vs_sequential']:.2f}, bitwise "
            f"{ae['sampled_gradients_bitwise_equal_to_single_runs']}")
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        out["cpu_baseline"] = cpu_baseline(mesh, u, h, tr, h_s, args.dt, args.tracers,
                                           steps_1=args.cpu_steps, steps_all=2 * args.cpu_steps)
        cb = out["cpu_baseline"]
        log(f"  cpu: {cb['value']:.3e} cell-steps/s on {cb['cores']} threads, "
            f"{cb['one_core']['value']:.3e} on one ({cb['cpu_model']}, nproc {cb['nproc']})")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
