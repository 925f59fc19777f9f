"""Non-finite values (NaN, Inf) in a state, forced from the host.

The rule: the uploads refuse them with MPAS_ENONFINITE (naming the field and
index), as the reference's HDF5 reads abort a conversion of NaN or Inf
(mpasoflecsi/specialization/definition.hh:165-191); the kernels do not test
values, so NaN propagates through the tasks as through the reference's loops;
mpas_sw_state_check counts the non-finite values of the resident state."""
import numpy as np
import pytest

from conftest import mesh

pytestmark = pytest.mark.gpu

ENONFINITE = -6


def _state(m, nT=2, seed=2):
    from mpas_hip import sw_workload_state
    u, h, tr = sw_workload_state(m, nT, seed=seed)
    return u, h, tr.ravel(), np.full(m.nCells, 4500.0)


def _ctx(m, nT=2):
    from mpas_hip import TRACER_FCT, Context
    return Context(m, nTracers=nT, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)


@pytest.mark.parametrize("field,index,value", [("h", 5, np.nan), ("u", 17, np.inf),
                                               ("tracers", 31, -np.inf), ("h_s", 0, np.nan)])
def test_upload_refuses_non_finite(mesh3, field, index, value):
    from mpas_hip import MpasError
    m = mesh3
    u, h, tr, h_s = _state(m)
    arrs = {"h": h, "u": u, "tracers": tr, "h_s": h_s}
    arrs[field] = arrs[field].copy()
    arrs[field][index] = value
    ctx = _ctx(m)
    print(f"upload with {value} in {field}[{index}]", flush=True)
    with pytest.raises(MpasError) as e:
        ctx.sw_upload(arrs["h"], arrs["u"], arrs["tracers"], arrs["h_s"])
    assert e.value.code == ENONFINITE
    assert f"{field}[{index}]" in str(e.value)
    # nothing was uploaded: stepping still needs a state
    with pytest.raises(MpasError) as e:
        ctx.rk4_step(600.0, 1)
    assert e.value.code == -4
    ctx.close()


def test_batch_upload_refuses_non_finite(mesh3):
    from mpas_hip import Batch, MpasError
    m = mesh3
    ctx = _ctx(m)
    b = Batch([ctx], [0, 0], [600.0, 600.0])
    u, h, tr, h_s = _state(m)
    b.upload(0, h, u, tr, h_s)
    u2 = u.copy()
    u2[3] = np.nan
    with pytest.raises(MpasError) as e:
        b.upload(1, h, u2, tr, h_s)
    assert e.value.code == ENONFINITE and "u[3]" in str(e.value)
    b.close()
    ctx.close()


def test_nan_propagates_through_the_diagnostics(mesh3):
    """compute_solve_diagnostics with NaN in h at one cell: h_edge is NaN on
    exactly that cell's edges and finite elsewhere (h_edge = (h1 + h2) / 2,
    apps/sw/tasks/solve.cc:31-39); ke, which reads u only, stays finite."""
    m = mesh3
    C, E, V = m.nCells, m.nEdges, m.nVertices
    u, h, _, _ = _state(m, nT=0)
    c = 123
    h = h.copy()
    h[c] = np.nan
    from mpas_hip import Context
    ctx = Context(m)
    dh, du = ctx.to_device(h), ctx.to_device(u)
    outs = [ctx.alloc(n) for n in (E, V, V, V, C, E, V)]
    ctx.compute_solve_diagnostics(dh, du, *outs)
    h_edge, ke = outs[0].numpy(), outs[4].numpy()
    eoc = m["edgesOnCell"].reshape(-1, m.maxEdges)[c]
    mine = set(int(e) - 1 for e in eoc if e > 0)
    bad = set(np.flatnonzero(~np.isfinite(h_edge)).tolist())
    assert bad == mine
    assert np.isfinite(ke).all()
    ctx.close()


def test_state_check_counts_blow_up(mesh3):
    """A finite run checks clean; a time step far beyond the gravity-wave
    limit (CFL about 25) overflows within tens of steps, and the check counts
    the non-finite values the step left (the step itself does not test)."""
    m = mesh3
    u, h, tr, h_s = _state(m)
    ctx = _ctx(m)
    ctx.sw_upload(h, u, tr, h_s)
    ctx.rk4_step(600.0, 3)
    assert ctx.state_check() == 0
    ctx.rk4_step(1e5, 200)
    n = ctx.state_check()
    print("non-finite values after the blow-up:", n, flush=True)
    h1, u1, t1 = ctx.sw_download()
    assert n == int((~np.isfinite(h1)).sum() + (~np.isfinite(u1)).sum() +
                    (~np.isfinite(t1)).sum())
    assert n > 0
    ctx.close()
