"""Pins for the native-tier contact law of the oracle (orc_native_contacts_f64),
the law that broken bonds and never-bonded pairs meet on the device: Hertz
closed form, the contact reach of a broken bond, viscous damping, the
tangential spring below and at the Coulomb cap, the turn of an old spring into
the tangent plane, and rolling resistance.  Each value below is worked out
from the formulas, not from the restatement."""
import math

import numpy as np
import pytest

from oracle.oracle import native_contacts

E, NU, RHO, E_REST = 30.0, 0.25, 2.65, 0.6
G = 0.5 * E / (1 + NU)
ESTAR = 0.5 * E / (1 - NU * NU)
GSTAR = 0.5 * G / (2 - NU)
_le = math.log(E_REST)
PA = _le / math.sqrt(math.pi ** 2 + _le * _le)
RHO43PI = RHO * 4.0 / 3.0 * math.pi
DT = 1e-3
RI, RJ = 0.4, 0.6
RS = RI * RJ / (RI + RJ)
MS = (RHO43PI * RI ** 3) * (RHO43PI * RJ ** 3) / (RHO43PI * (RI ** 3 + RJ ** 3))


def run(L, reach=RI + RJ, Vj=(0, 0, 0), Wi=(0, 0, 0), O=(0, 0, 0), mu=0.45, mur=0.05):
    R = np.array([[0.0, 0, 0], [L, 0, 0]])
    V = np.array([[0.0, 0, 0], Vj], np.float64)
    W = np.array([Wi, [0.0, 0, 0]], np.float64)
    Oa = np.array([O], np.float64)
    F, M = native_contacts(R, V, W, [RI, RJ], np.array([[0, 1]]), [reach], Oa, DT, ESTAR,
                           GSTAR, PA, mu, mur, RHO43PI)
    return F[0], M[0], Oa[0]


def hertz(delta):
    return 4.0 / 3.0 * ESTAR * math.sqrt(RS) * delta ** 1.5


def test_hertz_head_on():
    F, M, O = run(0.95)
    assert F[0] == pytest.approx(-hertz(0.05), rel=1e-13)
    assert F[1] == 0 and F[2] == 0
    assert np.all(M == 0) and np.all(O == 0)


def test_broken_bond_reach():
    # a pair bonded at L0 = 0.97 < ri + rj: no force at the rest length, and
    # the overlap counts from L0 once pushed closer
    F, _, _ = run(0.97, reach=0.97)
    assert np.all(F == 0)
    F, _, _ = run(0.95, reach=0.97)
    assert F[0] == pytest.approx(-hertz(0.02), rel=1e-13)


def test_beyond_reach_clears_spring():
    F, M, O = run(1.01, O=(0, 1e-3, 0))
    assert np.all(F == 0) and np.all(M == 0) and np.all(O == 0)


def test_normal_damping():
    v = 0.02  # j approaches i
    F, _, _ = run(0.95, Vj=(-v, 0, 0))
    Kn = 2 * ESTAR * math.sqrt(RS * 0.05)
    dmp = -1.8257 * PA * (-v) * math.sqrt(Kn * MS)
    assert dmp < 0  # approaching: more repulsion on i (towards -x)
    assert F[0] == pytest.approx(-hertz(0.05) + dmp, rel=1e-13)


def test_tangential_spring_below_cap():
    vy = 1e-3
    F, M, O = run(0.95, Vj=(0, vy, 0))
    Kt = 8 * GSTAR * math.sqrt(RS * 0.05)
    ft = Kt * vy * DT - 1.8257 * PA * math.sqrt(Kt * MS) * vy
    assert abs(ft) < 0.45 * hertz(0.05)
    assert F[1] == pytest.approx(ft, rel=1e-12)
    assert O[1] == pytest.approx(vy * DT, rel=1e-12) and O[0] == 0 and O[2] == 0
    # n x ft = x * (ft y) = ft z, lever ri
    assert M[2] == pytest.approx(ft * RI, rel=1e-12) and M[0] == 0 and M[1] == 0


def test_coulomb_cap():
    vy = 50.0
    F, _, O = run(0.95, Vj=(0, vy, 0))
    Kt = 8 * GSTAR * math.sqrt(RS * 0.05)
    assert F[1] == pytest.approx(0.45 * hertz(0.05), rel=1e-12)
    assert O[1] == pytest.approx(0.45 * hertz(0.05) / Kt, rel=1e-12)


def test_old_spring_turns_into_tangent_plane():
    # a spring with a normal part is turned into the plane, length kept
    a = 1e-6
    F, _, O = run(0.95, O=(a, a, 0))
    assert O[0] == 0 and O[1] == pytest.approx(math.sqrt(2) * a, rel=1e-12)
    Kt = 8 * GSTAR * math.sqrt(RS * 0.05)
    assert F[1] == pytest.approx(Kt * math.sqrt(2) * a, rel=1e-12)


def test_rolling_resistance():
    w = 0.3
    _, M1, _ = run(0.95, Wi=(0, 0, w))
    _, M0, _ = run(0.95, Wi=(0, 0, w), mur=0.0)
    # the spin of i also slides the contact (n x W ri = -w ri y), so compare
    # with and without rolling resistance: the difference is -mur |fn| W/|W|
    d = M1 - M0
    assert d[2] == pytest.approx(-0.05 * hertz(0.05), rel=1e-12)
    assert d[0] == 0 and d[1] == 0


def test_mirror_pair_balances():
    # i sees j and j sees i: equal and opposite loads with no velocities
    R = np.array([[0.0, 0, 0], [0.6, 0.5, 0.3]])
    L = float(np.linalg.norm(R[1]))
    Z = np.zeros((2, 3))
    Oa = np.zeros((2, 3))
    F, M = native_contacts(R, Z, Z, [RI, RJ], np.array([[0, 1], [1, 0]]), [L + 0.05] * 2, Oa,
                           DT, ESTAR, GSTAR, PA, 0.45, 0.05, RHO43PI)
    np.testing.assert_allclose(F[0], -F[1], rtol=1e-14)
    assert np.all(M == 0)
