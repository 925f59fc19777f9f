"""The run's output file (apps/sw/run.cc:73-113) on the host: which steps are
written (output_flagger over block_partition, mpasoflecsi/io/state.hh:22-38,
util/part.hh:28-53) and the file layout of mesh_output_init /
field_output_wrapper (tasks/io.cc:160-244, tasks/io.hh:13-47,
io/helpers.hh:127-258): datasets [Time][entity](level)(tracer) with their
dimension scales attached."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, bitwise_equal, mesh


def _io_built():
    from mpas_hip.meshfile import IO_LIB
    return os.path.exists(IO_LIB)


def _block_ends(n, P):
    """Independent statement of the rule: the partition splits 0..n-1 into P
    blocks starting at b*n//P; a step i < n is written when i+1 starts a
    block (i ends one).  Step n lies outside the partition: its owner is
    P + q with q = (P-1)//n, whose end n + (q+1)*n//P - 1 is n exactly when
    (q+1)*n < 2P (for P <= n: when n // P == 1)."""
    starts = {b * n // P for b in range(P + 1)}
    out = [i for i in range(1, n) if i + 1 in starts]
    if ((P - 1) // n + 1) * n < 2 * P:
        out.append(n)
    return out


def test_output_steps_worked_examples():
    from mpas_hip.output import output_steps
    # 10 steps, 5 output times: blocks {0,1} {2..4} {5,6} {7..9} -> 1, 4, 6, 9
    assert output_steps(10, 5) == [1, 4, 6, 9]
    # n a multiple of P: block ends k*n/P - 1; the final step is not written
    assert output_steps(100, 11) == [10 * k - 1 for k in range(1, 11)]
    # n == P: every block has one index; block 0 ends at 0 (never written)
    # and step n falls into block P, which ends at n
    assert output_steps(4, 5) == [1, 2, 3, 4]
    assert output_steps(3, 5) == [1, 2, 3]
    assert output_steps(0, 5) == [] and output_steps(10, 1) == [] and output_steps(10, 0) == []


def test_output_steps_against_block_ends():
    from mpas_hip.output import output_steps
    for n in range(1, 60):
        for T in range(2, 40):
            got = output_steps(n, T)
            assert got == _block_ends(n, T - 1), (n, T)
            assert len(got) <= T - 1  # rows 1..T-1 suffice: no write out of range
            assert all(1 <= a < b <= n for a, b in zip(got, got[1:]))


needs_io = pytest.mark.skipif(not _io_built(), reason="libmpas_io.so not built (no HDF5)")


@needs_io
def test_io_library_exports_and_binds_every_declared_symbol():
    from mpas_hip.meshfile import io
    from mpas_hip.output import _SIG, _io
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "mpas_io.h")).read(),
                 flags=re.S)
    names = sorted(set(re.findall(r"\b(mpas_[a-z0-9_]+)\s*\(", txt)))
    assert len(names) >= 12
    lib = _io()
    assert not [n for n in names if not hasattr(lib, n)]
    bound = {n for n in names if n in _SIG or getattr(io(), n).argtypes is not None}
    assert bound == set(names), set(names) - bound


@needs_io
def test_output_file_layout_and_round_trip(tmp_path):
    from mpas_hip.output import (CELLS, VERTICES, OutputFile, read_output, scale_name)
    m = mesh(3)
    L, nT, T = 2, 3, 4
    p = str(tmp_path / "output.ncdf")
    rng = np.random.default_rng(5)
    rows = []
    with OutputFile(p, m, T, nLevels=L, nTracers=nT) as f:
        for t in range(3):  # time 3 is never written
            r = dict(h=rng.random((m.nCells, L)), h_s=rng.random(m.nCells),
                     vorticity=rng.random((m.nVertices, L)),
                     potential_vorticity=rng.random((m.nVertices, L)),
                     tracers=rng.random((m.nCells, L, nT)))
            f.write(t, r["h"], r["h_s"], r["vorticity"], r["potential_vorticity"], r["tracers"])
            rows.append(r)
    shapes = {"h": (T, m.nCells, L), "h_s": (T, m.nCells), "vorticity": (T, m.nVertices, L),
              "potential_vorticity": (T, m.nVertices, L), "tracers": (T, m.nCells, L, nT)}
    for name, shape in shapes.items():
        got = read_output(p, name)
        assert got.shape == shape, name
        for t in range(3):
            assert bitwise_equal(got[t], rows[t][name]), (name, t)
        assert not got[3].any(), name
    # dimension scales, as tasks/io.hh:24-29 attaches them
    assert [scale_name(p, "tracers", d) for d in range(4)] == \
        ["Time", "nCells", "nVertLevels", "nTracers"]
    assert [scale_name(p, "potential_vorticity", d) for d in range(3)] == \
        ["Time", "nVertices", "nVertLevels"]
    assert [scale_name(p, "h_s", d) for d in range(2)] == ["Time", "nCells"]
    # the mesh fields of mesh_output_init
    for name, n in (("nCells", m.nCells), ("nEdges", m.nEdges), ("nVertices", m.nVertices),
                    ("vertexDegree", 3), ("nVertLevels", L), ("Time", T), ("nTracers", nT)):
        assert read_output(p, name).shape == (n,), name
    for name, space in (("areaCell", "nCells"), ("latVertex", "nVertices"), ("dvEdge", "nEdges")):
        assert bitwise_equal(read_output(p, name), m[name]), name
        assert scale_name(p, name, 0) == space
    assert bitwise_equal(read_output(p, "mDensity"), m["meshDensity"])
    cov = read_output(p, "cellsOnVertex")
    assert cov.shape == (m.nVertices, 3)
    assert np.array_equal(cov.ravel(), m["cellsOnVertex"].astype(np.float64))
    assert scale_name(p, "cellsOnVertex", 1) == "vertexDegree"


@needs_io
def test_output_file_errors(tmp_path):
    from mpas_hip import MpasError
    from mpas_hip.output import CELLS, VERTICES, OutputFile
    m = mesh(3)
    f = OutputFile(str(tmp_path / "o.ncdf"), m, 3, nLevels=1)
    h = np.ones(m.nCells)
    with pytest.raises(MpasError):
        f.write_field(3, "h", h, CELLS, 1)  # time outside [0, 3)
    with pytest.raises(MpasError):
        f.write_field(-1, "h", h, CELLS, 1)
    with pytest.raises(MpasError):
        f.write_field(0, "h", np.ones(2 * m.nCells), CELLS, 2)  # levels differ from the file
    f.write_field(0, "h", h, CELLS, 1)
    with pytest.raises(MpasError):
        f.write_field(1, "h", np.ones(m.nVertices), VERTICES, 1)  # shape of its first write
    f.close()
    f.close()  # idempotent
