"""Pins for the calibration outer loop's arithmetic (rfd/calibrate.py) on a
linear response model R(p) = R* + (p - p*) J, where finite differences are
exact and the weighted misfit is quadratic:
* the first step is steepest descent with the exact line search (found here by
  a scalar minimiser, not by the update's own formula);
* the second direction is Fletcher-Reeves;
* repeated steps drive the misfit to zero (5 parameters, 3 results);
* the reference's index quirk agrees with the default exactly when all
  parameter variations are equal, and not otherwise;
* the contractions hold the lower bounds, and the reference quirk sets the
  tensile strength from E.
"""
import numpy as np
import pytest
from scipy.optimize import minimize_scalar

from rfd import calibrate as cal

rng = np.random.default_rng(11)
J = rng.standard_normal((cal.NPARAM, cal.NRESULT))
P_STAR = np.array([3.0, 1.5, 2.0, 0.8, 0.6])
GOALS = np.array([0.9, 1.2, 0.7])


def model(p):
    return GOALS + (np.asarray(p) - P_STAR) @ J


def results_of(st):
    return np.array([model(q) for q in st.parameter_sets()])


def F(p, st):
    D = st.goals - model(p)
    return float((st.weights * D * D).sum())


def new_state(quirks=False):
    return cal.CGState(params=P_STAR * np.array([1.3, 0.8, 1.1, 1.2, 0.9]), goals=GOALS.copy(),
                       reference_quirks=quirks)


def test_first_step_is_exact_line_search():
    st = new_state()
    p0 = st.params.copy()
    g = -2.0 * J @ (st.weights * (GOALS - model(p0)))  # analytic gradient of F
    t = minimize_scalar(lambda t: F(p0 - t * g, st), bracket=(0, 1e-3), tol=1e-14).x
    p1 = st.update(results_of(st))
    np.testing.assert_allclose(p1, p0 - t * g, rtol=1e-6)
    assert F(p1, st) < F(p0, st)


def test_second_direction_is_fletcher_reeves():
    st = new_state()
    st.update(results_of(st))
    g0, d0 = st.grad.copy(), st.direction.copy()
    p1 = st.params.copy()
    g1 = -2.0 * J @ (st.weights * (GOALS - model(p1)))
    st.update(results_of(st))
    beta = (g1 @ g1) / (g0 @ g0)
    np.testing.assert_allclose(st.direction, g1 + beta * d0, rtol=1e-6)


def test_misfit_goes_to_zero():
    st = new_state()
    f0 = F(st.params, st)
    for _ in range(12):
        st.update(results_of(st))
    assert F(st.params, st) < 1e-8 * f0


def test_reference_index_quirk():
    # equal variations: the quirk's denominator is the same number
    a = cal.CGState(params=np.full(5, 2.0), goals=GOALS.copy())
    b = cal.CGState(params=np.full(5, 2.0), goals=GOALS.copy(), reference_quirks=True)
    R = np.array([model(q) for q in a.parameter_sets()])
    a.update(R)
    b.update(R.copy())
    # (the new parameters differ: the quirk's contraction rule moves Ec)
    np.testing.assert_array_equal(a.grad, b.grad)
    np.testing.assert_array_equal(a.direction, b.direction)
    c, d = new_state(), new_state(True)
    R = results_of(c)
    c.update(R)
    d.update(R.copy())
    assert not np.allclose(c.grad, d.grad)


def test_contractions():
    S, L = cal.STRESS, cal.LENGTH
    p = cal.contractions([1.0, 1e-9, 1e-9, 1e-12, 1e-12])
    assert p[0] == 1.0 and p[1] == pytest.approx(1e5 / S) and p[2] == pytest.approx(1e8 / S)
    assert p[3] == pytest.approx(1e-6 / L) and p[4] == pytest.approx(1e-5 / L)
    # the reference rule puts the tensile strength at 0.6 E for any rock
    q = cal.contractions(cal.GRANITE_START, reference_quirks=True)
    assert q[1] == pytest.approx(0.6 * cal.GRANITE_START[0])
    np.testing.assert_array_equal(cal.contractions(cal.GRANITE_START), cal.GRANITE_START)
