"""A mesh read from an MPAS mesh file drives the fused step to the same bits
as the generated mesh it was written from."""
import os

import numpy as np
import pytest

from conftest import bitwise_equal, mesh

pytestmark = pytest.mark.gpu


def test_file_mesh_fused_step_bitwise(tmp_path):
    from mpas_hip import TRACER_FCT, Context, sw_workload_state
    from mpas_hip.meshfile import IO_LIB, MeshFile, write_mesh_file
    if not os.path.exists(IO_LIB):
        pytest.skip("libmpas_io.so not built (no HDF5)")
    m = mesh(4)
    p = tmp_path / "L4.nc"
    write_mesh_file(p, m)
    f = MeshFile(p)
    u, h, tr = sw_workload_state(m, 2, seed=4)
    out = []
    for src in (m, f):
        c = Context(src, nTracers=2, tracer_scheme=TRACER_FCT, visc2=1e5, kappa2=1e3)
        c.sw_upload(h, u, tr.ravel(), np.full(m.nCells, 4500.0))
        c.rk4_step(600.0, 4)
        out.append(c.sw_download())
        c.close()
    for a, b in zip(*out):
        assert bitwise_equal(a, b)
