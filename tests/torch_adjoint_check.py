"""Adjoint of the fused RK4 step behind a torch.autograd.Function (test
harness only; run as its own process by tests/test_gpu_adjoint.py).

torch is imported before the library: torch ships its own HIP runtime under
the soname ROCm uses, and the first one loaded serves the whole process.
Prints one JSON line: the observation loss, and its directional derivative
from the adjoint gradient against the one from the tangent-linear model."""
import json
import os
import sys

import torch  # noqa: F401  (first: its HIP runtime serves the library too)
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from conftest import mesh  # noqa: E402


def main():
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(4)
    ctx = Context(m, nTracers=1, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    u, h, tr = sw_workload_state(m, 1, seed=1)
    tr = tr.ravel()
    ctx.sw_upload(h, u, tr, np.full(m.nCells, 4500.0))
    ctx.adj_init(4)
    dt, nsteps = 600.0, 8
    dev = torch.device("cuda:0")

    class AdjointRun(torch.autograd.Function):
        """nsteps fused RK4 steps; backward = the library's adjoint sweep."""

        @staticmethod
        def forward(fctx, u0, h0, tr0):
            u1, h1, tr1 = torch.empty_like(u0), torch.empty_like(h0), torch.empty_like(tr0)
            torch.cuda.synchronize()
            ctx.adj_forward(dt, nsteps, (u0.data_ptr(), h0.data_ptr(), tr0.data_ptr()),
                            (u1.data_ptr(), h1.data_ptr(), tr1.data_ptr()))
            ctx.sync()
            return u1, h1, tr1

        @staticmethod
        def backward(fctx, gu, gh, gtr):
            gu, gh, gtr = (g.contiguous() for g in (gu, gh, gtr))
            ou, oh, otr = torch.empty_like(gu), torch.empty_like(gh), torch.empty_like(gtr)
            torch.cuda.synchronize()
            ctx.adj_backward((gu.data_ptr(), gh.data_ptr(), gtr.data_ptr()),
                             (ou.data_ptr(), oh.data_ptr(), otr.data_ptr()))
            ctx.sync()
            return ou, oh, otr

    rng = np.random.default_rng(5)
    obs = torch.as_tensor(rng.choice(m.nCells, 100, replace=False), device=dev)
    u0, h0, tr0 = (torch.as_tensor(a, device=dev).requires_grad_(True) for a in (u, h, tr))
    with torch.no_grad():  # observations: a forward run from a perturbed state
        _, hp, trp = AdjointRun.apply(u0 + 0.1, h0 + 1.0, tr0 * 1.01)
        h_obs, tr_obs = hp[obs].clone(), trp[obs].clone()
    _, h1, tr1 = AdjointRun.apply(u0, h0, tr0)
    J = ((h1[obs] - h_obs) ** 2).sum() + ((tr1[obs] - tr_obs) ** 2).sum()
    J.backward()
    grads = [x.grad.cpu().numpy() for x in (u0, h0, tr0)]
    # the same directional derivative by the tangent-linear model
    dx_h = [rng.standard_normal(m.nEdges), rng.standard_normal(m.nCells),
            0.01 * rng.standard_normal(m.nCells)]
    x0, dx, xN, dxN = (ctx.state_arrays() for _ in range(4))
    for a, v in zip(x0, (u, h, tr)):
        a.upload(v)
    for a, v in zip(dx, dx_h):
        a.upload(v)
    ctx.tlm_forward(dt, nsteps, x0, dx, x1=xN, dx1=dxN)
    ctx.sync()
    o = obs.cpu().numpy()
    hN, tN = xN[1].numpy(), xN[2].numpy()
    dhN, dtN = dxN[1].numpy(), dxN[2].numpy()
    jt = float(np.sum(2 * (hN[o] - h_obs.cpu().numpy()) * dhN[o]) +
               np.sum(2 * (tN[o] - tr_obs.cpu().numpy()) * dtN[o]))
    ja = sum(float(np.dot(g, d)) for g, d in zip(grads, dx_h))
    ctx.close()
    print(json.dumps({"loss": J.item(), "adjoint": ja, "tlm": jt,
                      "rel_adjoint_vs_tlm": abs(ja - jt) / abs(jt),
                      "grad_finite": bool(all(np.isfinite(g).all() for g in grads)
                                          and np.abs(grads[1]).max() > 0)}), flush=True)


if __name__ == "__main__":
    main()
