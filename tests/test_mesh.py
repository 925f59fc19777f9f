"""Invariants of the generated meshes (the reference's own meshes are missing
blobs, so every workload runs on these)."""
import numpy as np
import pytest

from conftest import mesh


@pytest.mark.parametrize("level", [0, 1, 2, 3, 4])
def test_counts_and_euler(level):
    m = mesh(level)
    assert m.nCells == 10 * 4 ** level + 2
    assert m.nEdges == 30 * 4 ** level
    assert m.nVertices == 20 * 4 ** level
    assert m.nVertices - m.nEdges + m.nCells == 2


@pytest.mark.parametrize("level", [2, 4])
def test_areas_and_kites(level):
    m = mesh(level)
    a = m.radius
    sphere = 4 * np.pi * a * a
    assert abs(m["areaCell"].sum() / sphere - 1) < 1e-13
    assert abs(m["areaTriangle"].sum() / sphere - 1) < 1e-13
    kite = m["kiteAreasOnVertex"].reshape(-1, 3)
    assert np.allclose(kite.sum(1), m["areaTriangle"], rtol=1e-13, atol=0)
    cov = m["cellsOnVertex"].reshape(-1, 3) - 1
    acc = np.zeros(m.nCells)
    np.add.at(acc, cov.ravel(), kite.ravel())
    assert np.allclose(acc, m["areaCell"], rtol=1e-12, atol=0)
    assert (kite > 0).all()


@pytest.mark.parametrize("level", [2, 4])
def test_mpas_conventions(level):
    m = mesh(level)
    coe = m["cellsOnEdge"].reshape(-1, 2)
    assert (coe[:, 0] < coe[:, 1]).all(), "edges must point from the lower to the higher cell"
    cov = m["cellsOnVertex"].reshape(-1, 3)
    assert (np.diff(cov, axis=1) > 0).all(), "cellsOnVertex is written in ascending id"
    # cellsOnCell[j] is the cell across edgesOnCell[j]
    n = m["nEdgesOnCell"]
    eoc = m["edgesOnCell"].reshape(-1, 8)
    coc = m["cellsOnCell"].reshape(-1, 8)
    for c in range(m.nCells):
        for j in range(n[c]):
            e = eoc[c, j] - 1
            other = coe[e, 1] if coe[e, 0] - 1 == c else coe[e, 0]
            assert coc[c, j] == other


@pytest.mark.parametrize("level", [3, 4])
def test_trisk_weights_antisymmetric(level):
    """dc_e dv_e w(e,e') = -dc_e' dv_e' w(e',e) (Thuburn et al. 2009), the
    property that makes the PV flux in compute_tend energy neutral."""
    m = mesh(level)
    dc, dv = m["dcEdge"], m["dvEdge"]
    neoe = m["nEdgesOnEdge"]
    eoe = m["edgesOnEdge"].reshape(-1, 16) - 1
    w = m["weightsOnEdge"].reshape(-1, 16)
    W = {}
    for e in range(m.nEdges):
        for j in range(neoe[e]):
            W[(e, int(eoe[e, j]))] = dc[e] * dv[e] * w[e, j]
    scale = max(abs(x) for x in W.values())
    assert max(abs(W[(e, f)] + W[(f, e)]) for (e, f) in W) < 1e-13 * scale


def test_level8_sizes():
    m = mesh(8)
    assert (m.nCells, m.nEdges, m.nVertices) == (655362, 1966080, 1310720)
