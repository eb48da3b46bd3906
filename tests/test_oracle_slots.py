"""Pins of the oracle's superseded per-slot force path (orc_forces_slots:
md_calculateforcesdem.cu d_CalculateForcesDEM :328, _1 :501, _2 :581) and
penalty platens (orc_platen_uniaxial_penalty:
md_uniaxialcompression_simple.cu:29), against beam theory, Hertz-Mindlin
closed forms, hand sums, and the fused slot law on the same motion.

The per-slot path stores each slot's loads: Fijn = E A eps n, Fijt = G A O /
L0, Mijn = G J theta / L0 (J = pi b^4 / 2), Mijt = E I theta / L0
(I = pi b^4 / 4, without the extra bending factor of the fused kernel),
Mijadd = 0.5 r x Fijt for bonds; for contacts Mijt = rP n x Fijt and Mijadd
the rolling resistance.
"""
import dataclasses
import math

import numpy as np
import pytest

from oracle.oracle import lib, ptr
from rfd import cases
from rfd.params import Material

PI_F = np.float32(3.14159265358979)
SLOT = ("rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt", "Mijadd")


def mat():
    return Material(E=30.0, nu=0.25, rho=2.65, p_r=0.5, b_r=0.25, b_d=1.05, sc=1e9, tc=1e9)


class SlotPair:
    """Particle 0 at the origin and particle 1 at r, one slot each pointing
    at the other (slot 0 of row 0, slot K of row 1), plus K - 1 spare slots
    per row."""

    def __init__(self, r, bonded=True, L0=None, K=2):
        self.N, self.K = 2, K
        self.R = np.zeros(6, np.float32)
        self.R[1], self.R[3], self.R[5] = r
        self.V = np.zeros(6, np.float32)
        self.W = np.zeros(6, np.float32)
        S = 2 * K
        self.IL = np.full(S, 0xFFFFFFFF, np.uint32)
        self.IL[0], self.IL[K] = 1, 0
        self.ILtype = np.zeros(S, np.uint8)
        if bonded:
            self.ILtype[0] = self.ILtype[K] = 1
        L0 = L0 if L0 is not None else float(np.linalg.norm(np.asarray(r, np.float64)))
        self.iL = np.full(S, np.float32(1.0 / L0), np.float32)
        self.st = {k: np.zeros(3 * S, np.float32) for k in SLOT}
        self.F = np.zeros(6, np.float32)
        self.M = np.zeros(6, np.float32)

    def set_v(self, p, v):
        self.V[[p, p + 2, p + 4]] = v

    def set_w(self, p, w):
        self.W[[p, p + 2, p + 4]] = w

    def slot(self, key, s):
        return self.st[key][3 * s:3 * s + 3].astype(np.float64)

    def run(self, m, mode):
        E, G = (m.hm_E, m.hm_G) if mode == 1 else (m.E, m.G)
        lib().orc_forces_slots(mode, ptr(self.R), ptr(self.V), ptr(self.W), ptr(self.IL),
                               ptr(self.ILtype), ptr(self.iL), *(ptr(self.st[k]) for k in SLOT),
                               self.N, self.K, m.dt, E, G, m.b_r, m.p_A, m.p_mu, m.p_mur,
                               m.p_m, m.p_r)

    def sumup(self):
        lib().orc_sumup_forces(*(ptr(self.st[k]) for k in ("Fijn", "Fijt", "Mijn", "Mijt",
                                                           "Mijadd")),
                               ptr(self.F), ptr(self.M), self.N, self.K)


# ---------------------------------------------------------------- bond slots
@pytest.mark.parametrize("mode", [0, 2])
def test_axial_stretch(mode):
    m = mat()
    p = SlotPair((1.001, 0, 0), L0=1.0)
    p.run(m, mode)
    area = float(PI_F) * m.b_r ** 2
    f = p.slot("Fijn", 0)
    assert f[0] == pytest.approx(m.E * area * 1e-3, rel=2e-4)
    assert f[1] == f[2] == 0
    assert np.array_equal(p.slot("Fijn", p.K), -f)
    for k in ("Fijt", "Mijn", "Mijt", "Mijadd", "Oijt"):
        assert not p.st[k].any(), k
    np.testing.assert_array_equal(p.slot("rij", 0), [1, 0, 0])
    np.testing.assert_array_equal(p.slot("rij", p.K), [-1, 0, 0])


def test_shear_spring_force_and_moment():
    """j slides along +y at v for k steps: O = v dt k, Fijt = G A O / L0 on
    the slot of i, Mijadd = 0.5 L x Fijt (about +z)."""
    m = mat()
    p = SlotPair((1.0, 0, 0), L0=1.0)
    v, k = 1e-3, 20
    p.set_v(1, (0, v, 0))
    for _ in range(k):
        p.run(m, 2)
    delta = v * m.dt * k
    Ft = m.G * float(PI_F) * m.b_r ** 2 * delta
    assert p.slot("Oijt", 0)[1] == pytest.approx(delta, rel=1e-4)
    assert p.slot("Fijt", 0)[1] == pytest.approx(Ft, rel=1e-4)
    assert p.slot("Mijadd", 0)[2] == pytest.approx(0.5 * Ft, rel=1e-4)
    p.sumup()
    assert p.F[0 + 2] == pytest.approx(Ft, rel=1e-4)


def test_twist_moment_is_torsion():
    """j spins about the bond axis: Mijn = G J theta / L0, J = pi b^4 / 2."""
    m = mat()
    p = SlotPair((1.0, 0, 0), L0=1.0)
    w, k = 2e-3, 25
    p.set_w(1, (w, 0, 0))
    for _ in range(k):
        p.run(m, 2)
    theta = w * m.dt * k
    J = float(PI_F) * m.b_r ** 4 / 2.0
    mn = p.slot("Mijn", 0)
    assert mn[0] == pytest.approx(m.G * J * theta, rel=1e-4)
    assert p.slot("Mijn", p.K)[0] == pytest.approx(-mn[0], rel=1e-6)
    assert not p.st["Mijt"].any() and not p.st["Fijn"].any()


def test_bend_moment_is_beam_bending():
    """j rotates about +y: Mijt = E I theta / L0 with I = pi b^4 / 4 and no
    further factor (the fused kernel's extra dt E / (4 L0) is absent)."""
    m = mat()
    p = SlotPair((1.0, 0, 0), L0=1.0)
    w, k = 2e-3, 10
    p.set_w(1, (0, w, 0))
    for _ in range(k):
        p.run(m, 2)
    theta = w * m.dt * k
    I = float(PI_F) * m.b_r ** 4 / 4.0
    assert p.slot("Mijt", 0)[1] == pytest.approx(m.E * I * theta, rel=1e-4)
    # the rotation also shears the bond: O = 0.5 w L dt k along +z
    O = 0.5 * w * m.dt * k
    assert p.slot("Fijt", 0)[2] == pytest.approx(m.G * float(PI_F) * m.b_r ** 2 * O, rel=1e-4)


def test_rigid_translation_is_load_free():
    m = mat()
    p = SlotPair((0.6, 0.7, -0.2))
    p.set_v(0, (0.3, -0.1, 0.2))
    p.set_v(1, (0.3, -0.1, 0.2))
    for _ in range(5):
        p.run(m, 2)
    # the rest length is 1/iL rounded to float32: a stretch of an ulp or so
    assert np.abs(p.st["Fijn"]).max() < 1e-6 * m.E * float(PI_F) * m.b_r ** 2
    for k in ("Fijt", "Mijn", "Mijt", "Mijadd"):
        assert not p.st[k].any(), k


def test_modes_zero_and_leave_slots_as_the_reference_does():
    """Slot 1 of each row is empty, row 0 slot 0 a bond; a second pair of
    rows holds a contact slot beyond 2 rP.  mode 0 zeroes empty slots and
    leaves contact slots; mode 1 leaves bond slots and zeroes empty and
    far contact slots; mode 2 touches bond slots only."""
    m = mat()
    for mode in (0, 1, 2):
        p = SlotPair((1.001, 0, 0), L0=1.0)
        # row 1's slot K is turned into a contact slot beyond reach
        p.ILtype[p.K] = 0
        for k in SLOT:
            p.st[k][:] = 7.0
        p.run(m, mode)
        empty = [1, p.K + 1]
        for s in empty:
            for k in ("Fijn", "Fijt", "Oijt", "Mijn", "Mijt", "Mijadd"):
                want = 7.0 if mode == 2 else 0.0
                assert np.all(p.slot(k, s) == want), (mode, k, s)
        for k in ("Fijn", "Fijt", "Oijt", "Mijn", "Mijt", "Mijadd"):
            # the contact slot (1.001 apart, beyond 2 rP = 1.0)
            want = 0.0 if mode == 1 else 7.0
            assert np.all(p.slot(k, p.K) == want), (mode, k)
            # the bond slot
            changed = not np.all(p.slot(k, 0) == 7.0)
            assert changed == (mode != 1), (mode, k)


# ------------------------------------------------------------- contact slots
def test_hertz_normal_force():
    m = mat()
    delta = 0.01
    p = SlotPair((2 * m.p_r - delta, 0, 0), bonded=False)
    p.run(m, 1)
    want = -(4.0 / 3.0) * m.hm_E * math.sqrt(m.p_r / 2.0) * delta ** 1.5
    assert p.slot("Fijn", 0)[0] == pytest.approx(want, rel=1e-4)
    assert np.array_equal(p.slot("Fijn", p.K), -p.slot("Fijn", 0))
    assert not p.st["Mijn"].any()


def test_coulomb_cap_and_tangential_moment():
    """A large tangential spring is capped at mu |fn|; the spring is reset to
    Fijt / Kt; Mijt = rP n x Fijt."""
    m = mat()
    delta = 0.01
    p = SlotPair((2 * m.p_r - delta, 0, 0), bonded=False)
    p.st["Oijt"][0:3] = [0, 0.05, 0]
    p.run(m, 1)
    fn = abs(p.slot("Fijn", 0)[0])
    ft = p.slot("Fijt", 0)
    assert np.linalg.norm(ft) == pytest.approx(m.p_mu * fn, rel=1e-5)
    Kt = 8.0 * m.hm_G * math.sqrt(m.p_r * delta / 2.0)
    np.testing.assert_allclose(p.slot("Oijt", 0), ft / Kt, rtol=1e-4, atol=1e-12)
    np.testing.assert_allclose(p.slot("Mijt", 0), m.p_r * np.cross([1, 0, 0], ft), rtol=1e-5,
                               atol=1e-12)


def test_rolling_resistance():
    m = mat()
    delta = 0.01
    p = SlotPair((2 * m.p_r - delta, 0, 0), bonded=False)
    p.set_w(0, (0, 3.0, 4.0))
    p.run(m, 1)
    fn = abs(p.slot("Fijn", 0)[0])
    np.testing.assert_allclose(p.slot("Mijadd", 0), -m.p_mur * fn * np.array([0, 0.6, 0.8]),
                               rtol=1e-5)
    # particle 1 does not spin: no rolling moment on its slot
    assert not p.slot("Mijadd", p.K).any()


# --------------------------------------------- against the fused slot law
def test_per_slot_bonds_match_fused_law_on_a_moving_lattice():
    """Bonds of a lattice cylinder under random velocities and spins, five
    steps: the per-slot path's summed forces equal the fused kernel's within
    FP32 rounding, and its stored moments are the fused kernel's stored
    twist and bending states times pi b^4 (both accumulate the same
    rotation sequence)."""
    from oracle.oracle import OracleLayout

    case = cases.reference_uniaxial(5.0, 6.0, seed=4)
    m = case.mat
    aacut = float(np.float32(m.a_farcut) ** 2)
    a, b = OracleLayout(case.R, case.K, case.grid()), OracleLayout(case.R, case.K, case.grid())
    g = np.random.default_rng(5)
    V = (g.standard_normal(3 * case.N) * 0.2).astype(np.float32)
    W = (g.standard_normal(3 * case.N) * 0.5).astype(np.float32)
    for o in (a, b):
        o.cell_distribution()
        o.il_construct(aacut)
        o.bonds_construct(m.b_dd)
        o.V[:] = V
        o.W[:] = W
    unbreakable = dataclasses.replace(m, sc=1e30, tc=1e30)
    for _ in range(5):
        a.F[:] = 0
        a.M[:] = 0
        a.forces_bpm(unbreakable)
        b.forces_slots(2, m)
        b.sumup_forces()
        for o in (a, b):
            o.R += o.V * np.float32(1e-3)
    assert (a.ILtype == 1).sum() > 4 * case.N
    scale = np.abs(a.F).max()
    np.testing.assert_allclose(b.F, a.F, rtol=0, atol=2e-5 * scale)
    j4 = float(PI_F) * m.b_r ** 4
    bond = np.repeat(a.ILtype == 1, 3)
    for k in ("Mijn", "Mijt"):
        x, y = b.slot[k][bond], a.slot[k][bond] * j4
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-5 * np.abs(y).max())


# ---------------------------------------------------------- penalty platens
def test_penalty_platens_hand_case():
    """Platens at z = 5 +- 4.5 with stiffness C = 100 and mu = 0.5:
    particle 0 sits 0.1 above the top (F.z -= 10, and friction 0.5 * 10 against
    its in-plane velocity (3, 4)/5), particle 1 0.2 below the bottom (F.z +=
    20), particle 2 between the platens, particle 3 outside the footprint,
    particle 4 above Zcut."""
    from oracle.oracle import OracleLayout
    from rfd.reflayout import Grid

    P = np.array([[0, 0, 9.6], [0.5, 0, 0.3], [0, 0, 5.0], [9, 9, 9.9], [0, 0, 12.0]],
                 np.float32)
    o = OracleLayout(np.ascontiguousarray(P.T).reshape(-1), 1, Grid(1, 1, 1, 1, 1))
    N = 5
    o.V[0], o.V[N] = 3.0, 4.0
    o.F[:] = 1.0
    sb, st = o.platen_uniaxial_penalty((0.0, 0.0, 5.0), 4.0, 4.5, 100.0, 0.5, 11.0)
    fz_top = (4.5 - (9.6 - 5.0)) * 100.0        # -10 (float32 arithmetic)
    fz_bot = (4.5 + (0.3 - 5.0)) * 100.0        # -20
    assert st == pytest.approx(-fz_top, rel=1e-5) and sb == pytest.approx(fz_bot, rel=1e-5)
    assert o.F[0 + 2 * N] == pytest.approx(1.0 + fz_top, rel=1e-5)
    assert o.F[0] == pytest.approx(1.0 - 0.5 * 10.0 * 0.6, rel=1e-5)
    assert o.F[0 + N] == pytest.approx(1.0 - 0.5 * 10.0 * 0.8, rel=1e-5)
    assert o.F[1 + 2 * N] == pytest.approx(1.0 - fz_bot, rel=1e-5)
    assert o.F[1] == 1.0 and o.F[1 + N] == 1.0       # at rest: no friction
    for i in (2, 3, 4):
        assert all(o.F[i + d * N] == 1.0 for d in range(3)), i


def test_moment_arms_in_a_general_direction():
    """Bond and contact along a general direction with general velocities and
    spins: every component of Mijadd = 0.5 r x Oijt G A / L0 (bond) and of
    Mijt = rP n x Fijt (contact), from the slot's own outputs."""
    m = mat()
    d = np.array([0.6, 0.7, -0.2])
    d /= np.linalg.norm(d)
    g = np.random.default_rng(3)
    for bonded, L in ((True, 1.0), (False, 2 * m.p_r - 0.01)):
        p = SlotPair(tuple(d * L), bonded=bonded)
        p.set_v(1, g.standard_normal(3) * 1e-2)
        p.set_w(0, g.standard_normal(3) * 1e-2)
        p.set_w(1, g.standard_normal(3) * 1e-2)
        p.st["Oijt"][0:3] = np.cross(d, g.standard_normal(3)) * 1e-4
        p.run(m, 2 if bonded else 1)
        r = p.R[[1, 3, 5]].astype(np.float64)
        if bonded:
            ct = m.G * float(PI_F) * m.b_r ** 2 * float(p.iL[0])
            want = 0.5 * np.cross(r, p.slot("Oijt", 0)) * ct
            got = p.slot("Mijadd", 0)
        else:
            n = r / np.linalg.norm(r)
            want = m.p_r * np.cross(n, p.slot("Fijt", 0))
            got = p.slot("Mijt", 0)
        assert np.abs(want).min() > 0
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
