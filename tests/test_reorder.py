"""Mesh reordering on the host (mpas_host_mesh_reorder): the renumbered mesh
is the original one seen through the permutations.  Every connectivity list
keeps its entries' order with the ids renamed, every per-entity array moves
with its entity, edges follow their new cell pair and vertices their sorted
new cell triple."""
import numpy as np
import pytest

from conftest import mesh

LISTS = {"c2v": ("c", "v"), "c2e": ("c", "e"), "c2c": ("c", "c"), "e2v": ("e", "v"),
         "e2e": ("e", "e"), "e2c": ("e", "c"), "v2c": ("v", "c"), "v2e": ("v", "e")}


def _rows(cnx):
    off, idx = cnx
    return [idx[off[i]:off[i + 1]] for i in range(len(off) - 1)]


@pytest.mark.parametrize("order", [0, 1, 2])
def test_reordered_mesh_is_a_renumbering(order):
    from mpas_hip import TRACER_FCT, HostMesh
    m = mesh(3)
    V = m.nVertices
    hm = HostMesh(m, nTracers=2, tracer_scheme=TRACER_FCT, kappa2=1e3)
    r = hm.reordered(order)
    of = dict(zip("cev", r.order(V)))
    new = {}
    for k, n in (("c", m.nCells), ("e", m.nEdges), ("v", V)):
        assert np.array_equal(np.sort(of[k]), np.arange(n)), k  # a permutation
        new[k] = np.empty(n, dtype=np.int64)
        new[k][of[k]] = np.arange(n)
    if order == 0:
        assert all(np.array_equal(of[k], np.arange(len(of[k]))) for k in of)
    for w, (src, dst) in LISTS.items():
        a, b = _rows(hm.connectivity(w)), _rows(r.connectivity(w))
        for i, row in enumerate(b):
            assert np.array_equal(row, new[dst][a[of[src][i]]]), (w, i)
    # edges by their (lower, higher) new cell pair
    e2c = np.array(_rows(r.connectivity("e2c")))
    key = np.sort(e2c, axis=1)
    assert np.all(np.diff(key[:, 0]) >= 0)
    # per-edge setup products move with their edge; stencil ids are renamed
    for f in ("advCoef", "advCoef_3rd", "derivTwo", "meshScalingDel2", "nAdvCellsForEdge"):
        a, b = hm.setup_field(f), r.setup_field(f)
        w = a.size // m.nEdges
        assert np.array_equal(b.reshape(m.nEdges, w), a.reshape(m.nEdges, w)[of["e"]]), f
    n = r.setup_field("nAdvCellsForEdge")
    a = hm.setup_field("advCellsForEdge").reshape(m.nEdges, 16)[of["e"]]
    b = r.setup_field("advCellsForEdge").reshape(m.nEdges, 16)
    for e in range(m.nEdges):
        assert np.array_equal(b[e, :n[e]], new["c"][a[e, :n[e]]])
    a, b = hm.setup_field("edgeSignOnCell"), r.setup_field("edgeSignOnCell")
    assert np.array_equal(b.reshape(m.nCells, 8), a.reshape(m.nCells, 8)[of["c"]])


def test_locality_orders_differ_from_the_given_one():
    from mpas_hip import HostMesh
    m = mesh(4)
    hm = HostMesh(m)
    cs = [hm.reordered(o).order(m.nVertices)[0] for o in (1, 2)]
    assert not np.array_equal(cs[0], np.arange(m.nCells))
    assert not np.array_equal(cs[1], np.arange(m.nCells))
    assert not np.array_equal(cs[0], cs[1])
