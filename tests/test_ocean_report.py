"""The ocean test cases' end-of-run error report (calc_error_case_1/2/3,
apps/ocean/tasks/testcase.cc:175-294) on the host: zero for the exact
solution, the expected norms for a uniform relative error, the
reference's max-error update rule, and the analytic solutions it scores
against (rotation by 2 pi elapsed / 1 day, exp(-6 kappa t / a^2) decay)."""
import numpy as np
import pytest

from conftest import mesh

A = 6371000.0


@pytest.mark.parametrize("which", [1, 2, 3])
def test_exact_solution_scores_zero(which):
    m = mesh(3)
    t = 3 * 86400.0 / 7
    _, _, truth = m.ocean_case(which, a=A, coef=1e7, elapsed=t)
    assert m.ocean_error(which, truth, t, a=A, coef=1e7) == (0.0, 0.0)


def test_uniform_relative_error():
    m = mesh(3)
    eps = 1e-3
    _, _, truth = m.ocean_case(1, a=A, elapsed=0.0)
    l2, mx = m.ocean_error(1, truth * (1 + eps), 0.0, a=A)
    # every cell's relative error is eps (to rounding): sqrt(n eps^2) / n
    assert l2 == pytest.approx(eps / np.sqrt(m.nCells), rel=1e-9)
    assert mx == pytest.approx(eps, rel=1e-9)


def test_max_error_follows_the_reference_update_rule():
    """errMax takes error whenever |error| > errMax: after a negative error
    it is overtaken by any later error, so the report can miss the largest
    one (testcase.cc:203-205 prints |errMax|)."""
    m = mesh(3)
    _, _, truth = m.ocean_case(1, a=A, elapsed=0.0)
    tr = truth.copy()
    tr[0] = truth[0] * (1 - 0.3)   # errMax = -0.3
    tr[5] = truth[5] * (1 + 1e-3)  # cells 1-4 (error 0 > -0.3) reset it to 0, then 1e-3
    l2, mx = m.ocean_error(1, tr, 0.0, a=A)
    assert mx == pytest.approx(1e-3, rel=1e-9)
    assert l2 == pytest.approx(np.sqrt(0.3 ** 2 + 1e-6) / m.nCells, rel=1e-9)


def test_stride_reads_tracer_zero():
    m = mesh(3)
    _, _, truth = m.ocean_case(3, a=A, coef=1e7, elapsed=0.0)
    tr = np.stack([truth, truth * 2, truth * 3], axis=1).ravel()  # [cell][tracer]
    assert m.ocean_error(3, tr, 0.0, a=A, coef=1e7, stride=3) == (0.0, 0.0)
    with pytest.raises(ValueError):
        m.ocean_error(3, truth[:10], 0.0, a=A)


def test_analytic_solutions():
    m = mesh(3)
    x, y = m["xCell"], m["yCell"]
    # case 1: a Gaussian in x rotated about the z axis; a full day returns it
    _, _, g0 = m.ocean_case(1, a=A, elapsed=0.0)
    _, _, g1 = m.ocean_case(1, a=A, elapsed=86400.0)
    assert np.allclose(g0, g1, rtol=1e-12)
    # a quarter day: x' = x cos(-pi/2) - y sin(-pi/2) = y (the hill moved east)
    _, _, gq = m.ocean_case(1, a=A, elapsed=21600.0)
    assert np.allclose(gq, 50.0 + 20.0 / 21983.7698959881 * np.exp(10.0 * y / A), rtol=1e-12)
    # case 2: the slotted cylinder takes the values 50 and 100 only
    _, _, s0 = m.ocean_case(2, a=A, elapsed=0.0)
    assert set(np.unique(s0)) == {50.0, 100.0}
    # case 3: the harmonic decays as exp(-6 kappa t / a^2)
    kappa, t = 1e7, 86400.0
    _, _, h0 = m.ocean_case(3, a=A, coef=kappa, elapsed=0.0)
    _, _, h1 = m.ocean_case(3, a=A, coef=kappa, elapsed=t)
    assert np.allclose(h1 - 50.0, (h0 - 50.0) * np.exp(-6 * kappa * t / A ** 2), rtol=1e-12,
                       atol=1e-12)
    assert x.size == m.nCells
