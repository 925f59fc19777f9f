"""MPAS mesh files (lib/libmpas_io.so): a generated mesh written and read
back field for field, and the flattened topology and FCT set-up of the read
mesh equal to the generated one's (the file path keeps the reference's
conventions: 1-based ids, zeros dropped anywhere in a row)."""
import os

import numpy as np
import pytest

from conftest import bitwise_equal, mesh

pytest.importorskip("mpas_hip.meshfile")


def _io_built():
    from mpas_hip.meshfile import IO_LIB
    return os.path.exists(IO_LIB)


pytestmark = pytest.mark.skipif(not _io_built(), reason="libmpas_io.so not built (no HDF5)")


def test_round_trip_every_field(tmp_path):
    from mpas_hip.mesh import DBL_FIELDS, INT_FIELDS
    from mpas_hip.meshfile import MeshFile, write_mesh_file
    m = mesh(3)
    p = tmp_path / "L3.nc"
    write_mesh_file(p, m)
    f = MeshFile(p)
    assert (f.nCells, f.nEdges, f.nVertices) == (m.nCells, m.nEdges, m.nVertices)
    assert (f.maxEdges, f.maxEdges2, f.vertexDegree) == (m.maxEdges, m.maxEdges2, 3)
    assert f.radius == m.radius
    for name in INT_FIELDS + DBL_FIELDS:
        assert bitwise_equal(f[name], m[name]), name
    f.close()


def test_read_mesh_flattens_like_generated(tmp_path):
    from mpas_hip import TRACER_FCT, HostMesh
    from mpas_hip.meshfile import MeshFile, write_mesh_file
    m = mesh(3)
    p = tmp_path / "L3.nc"
    write_mesh_file(p, m)
    f = MeshFile(p)
    a = HostMesh(m, nTracers=1, tracer_scheme=TRACER_FCT, kappa2=1e3)
    b = HostMesh(f, nTracers=1, tracer_scheme=TRACER_FCT, kappa2=1e3)
    for w in ["c2v", "c2e", "c2c", "e2v", "e2e", "v2c", "e2c", "v2e"]:
        x, y = a.connectivity(w), b.connectivity(w)
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]), w
    for name in ["advCellsForEdge", "advCoef", "advCoef_3rd", "meshScalingDel2"]:
        assert bitwise_equal(a.setup_field(name), b.setup_field(name)), name


def test_zero_padding_anywhere_in_a_row(tmp_path):
    """The padding zeros of the pentagons' rows moved to the front of the row:
    the flattened lists do not change (definition.hh:670-718 drops zeros
    wherever they sit)."""
    from mpas_hip import HostMesh
    from mpas_hip.meshfile import MeshFile, write_mesh_file

    class Shifted:
        def __init__(self, m):
            self.m = m
            self.nCells, self.nEdges, self.nVertices = m.nCells, m.nEdges, m.nVertices
            self.maxEdges, self.maxEdges2, self.radius = m.maxEdges, m.maxEdges2, m.radius

        def __getitem__(self, name):
            a = self.m[name]
            if name in ("edgesOnCell", "verticesOnCell", "cellsOnCell", "edgesOnEdge"):
                w = self.maxEdges2 if name == "edgesOnEdge" else self.maxEdges
                r = a.reshape(-1, w).copy()
                for i in range(r.shape[0]):
                    nz = r[i][r[i] != 0]
                    r[i] = np.concatenate([np.zeros(w - nz.size, r.dtype), nz])
                return r.ravel()
            return a

    m = mesh(2)
    sh = Shifted(m)
    assert (sh["edgesOnCell"].reshape(-1, m.maxEdges)[:, 0] == 0).any()  # pentagons moved
    p = tmp_path / "shifted.nc"
    write_mesh_file(p, sh)
    a, b = HostMesh(m), HostMesh(MeshFile(p))
    for w in ["c2v", "c2e", "c2c", "e2v", "e2e", "v2c", "e2c", "v2e"]:
        x, y = a.connectivity(w), b.connectivity(w)
        assert np.array_equal(x[1], y[1]), w


def test_file_errors(tmp_path):
    from mpas_hip import MpasError
    from mpas_hip.meshfile import MeshFile, write_mesh_file
    p = tmp_path / "text.nc"
    p.write_text("not hdf5")
    with pytest.raises(MpasError):
        MeshFile(p)
    m = mesh(2)
    q = tmp_path / "partial.nc"
    write_mesh_file(q, m, fields=["edgesOnCell", "areaCell"])
    with pytest.raises(MpasError):
        MeshFile(q)  # the connectivity and metrics are incomplete
