"""MI355X shallow-water RK4 + FCT step of MPAS-O-FleCSI behind a C ABI.

    from mpas_hip import Mesh, Context
    mesh = Mesh(level=4)
    ctx = Context(mesh, nTracers=4, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
    ctx.sw_upload(h, u, tracers, h_s)
    ctx.rk4_step(dt=960.0, nsteps=100)
"""
from ._lib import (BATCH_GRAPH, BATCH_LAUNCH, BATCH_PERSISTENT, BATCH_TUNED, ORDER_GIVEN,
                   ORDER_MORTON, ORDER_RCM, TRACER_CENTRED, TRACER_FCT,
                   TRACER_NONE, HIP_LIB, MESH_LIB, MpasError, hip)
from .mesh import EARTH_RADIUS, GRAVITY, OMEGA_F, REF_A, Mesh, sw_workload_state
from .model import Batch, Context, DeviceArray, HostMesh

__all__ = ["Batch", "BATCH_GRAPH", "BATCH_LAUNCH", "BATCH_PERSISTENT", "BATCH_TUNED", "ORDER_GIVEN", "ORDER_RCM", "ORDER_MORTON", "Context", "DeviceArray", "HostMesh", "Mesh", "MpasError", "sw_workload_state", "hip",
           "TRACER_NONE", "TRACER_CENTRED", "TRACER_FCT", "EARTH_RADIUS", "GRAVITY", "OMEGA_F",
           "REF_A", "HIP_LIB", "MESH_LIB"]
