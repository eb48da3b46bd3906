"""Pins for the packer's iteration in the oracle (oracle.pack_iteration):
two overlapping spheres separate by beta times their overlap along the line of
centres; a chain of three moves the middle sphere by the Jacobi sum; the wall
clamp puts a sphere's surface on the cylinder and the end planes; spheres that
do not overlap stay where they are."""
import math

import numpy as np
import pytest

from oracle.oracle import pack_iteration

BIG = dict(cx=0.0, cy=0.0, rc=100.0, z0=-100.0, z1=100.0)


def test_pair_separates_by_beta_times_overlap():
    R = np.array([[0.0, 0, 0], [0.9, 0, 0]])
    out, mo = pack_iteration(R, [0.5, 0.5], 0.5, **BIG)
    ov = 0.1
    assert out[0, 0] == pytest.approx(-0.25 * ov, abs=1e-8)
    assert out[1, 0] == pytest.approx(0.9 + 0.25 * ov, abs=1e-8)
    assert np.all(out[:, 1:] == 0)
    assert mo == pytest.approx(ov / 0.5)
    # with beta = 1 the pair ends exactly touching (to float32 rounding)
    out, _ = pack_iteration(R, [0.5, 0.5], 1.0, **BIG)
    assert out[1, 0] - out[0, 0] == pytest.approx(1.0, abs=1e-7)


def test_chain_jacobi_sum():
    R = np.array([[0.0, 0, 0], [0.9, 0, 0], [1.7, 0, 0]])
    out, _ = pack_iteration(R, [0.5, 0.5, 0.5], 0.5, **BIG)
    # middle: pushed +x by 0.25*0.1 from the left and -x by 0.25*0.2 from the right
    assert out[1, 0] == pytest.approx(0.9 + 0.025 - 0.05, abs=1e-7)
    assert out[0, 0] == pytest.approx(-0.025, abs=1e-7)
    assert out[2, 0] == pytest.approx(1.7 + 0.05, abs=1e-7)


def test_wall_and_end_clamps():
    R = np.array([[3.0, 4.0, 9.8], [0.0, 0.0, -9.9]])
    out, mo = pack_iteration(R, [0.5, 0.4], 0.5, cx=0.0, cy=0.0, rc=5.0, z0=-10.0, z1=10.0)
    assert mo == 0.0
    # first: radial distance 5 -> 4.5 (surface on the wall), direction kept
    assert math.hypot(out[0, 0], out[0, 1]) == pytest.approx(4.5, rel=1e-14)
    assert out[0, 0] / out[0, 1] == pytest.approx(0.75, rel=1e-14)
    assert out[0, 2] == pytest.approx(9.5)
    assert out[1, 2] == pytest.approx(-9.6)


def test_separated_spheres_stay():
    g = np.random.default_rng(1)
    R = g.uniform(-5, 5, (30, 3)) * np.array([1, 1, 1])
    R = R[np.argsort(R[:, 0])]
    rad = np.full(30, 0.01)
    out, mo = pack_iteration(R, rad, 0.5, **BIG)
    assert mo == 0.0
    np.testing.assert_array_equal(out, R)
