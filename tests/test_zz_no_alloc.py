"""mpas_rk4_step_noalloc (the fused step with its intermediates in the
library's fixed device scratch array) against mpas_rk4_step on an uploaded
state, through the C ABI: bit for bit, on two calls in a row, the second on
a smaller mesh, whose intermediates overwrite only part of what the first
call left in the array."""
import numpy as np
import pytest

from conftest import bitwise_equal, mesh

pytestmark = pytest.mark.gpu


def _case(level, nT, seed, dt, steps):
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    m = mesh(level)
    u, h, tr = sw_workload_state(m, nT, seed=seed)
    h_s = np.full(m.nCells, 4500.0)
    ctx = Context(m, nTracers=nT, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    ctx.sw_upload(h, u, tr, h_s)  # the existing entry point
    ctx.rk4_step(dt, steps)
    ref = ctx.sw_download()
    dh, du, dtr, dhs = (ctx.to_device(x) for x in (h, u, tr.ravel(), h_s))
    print(f"noalloc: level {level}, {nT} tracers, {steps} steps", flush=True)
    ctx.rk4_step_noalloc(dt, steps, dh, du, dtr, dhs)
    ctx.sync()
    got = (dh.numpy(), du.numpy(), dtr.numpy())
    ctx.close()
    return got, ref


def test_noalloc_equals_rk4_step_twice_then_smaller():
    for level, nT, seed, dt, steps in ((5, 4, 3, 480.0, 4), (3, 2, 8, 600.0, 3)):
        got, ref = _case(level, nT, seed, dt, steps)
        for g, r, name in zip(got, ref, ("h", "u", "tracers")):
            assert bitwise_equal(g, r.ravel()), (level, name)


def test_noalloc_rejects_a_step_beyond_the_scratch():
    from mpas_hip import TRACER_FCT, Context, MpasError
    ctx = Context(mesh(6), nLevels=2, nTracers=8, tracer_scheme=TRACER_FCT)
    n = ctx.alloc(1)
    with pytest.raises(MpasError) as e:
        ctx.rk4_step_noalloc(60.0, 1, n, n, n, n)
    assert e.value.code == -3 and "scratch" in str(e.value)
    ctx.close()
