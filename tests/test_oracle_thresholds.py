"""Pins for the parts of the CPU oracle that the closed-form tests of
test_oracle_pins.py leave open: the moment terms of both break criteria, the
velocity reset of the reference FIRE step, the full 8-cell slot order of the
list builders, the parameter derivation of potential_data, the tangential
part of the slot contact law, the platen clamps and the far-link prune.

Every expected value comes from mechanics or from hand arithmetic, not from
the oracle's own expressions:
  * bending of a beam of length L0 by a relative end rotation theta carries
    the surface stress E theta b / L0 (M = E I theta / L0, sigma = M b / I);
  * twisting it by theta carries the surface shear stress G theta b / L0
    (T = G J theta / L0, tau = T b / J);
  * sliding one end sideways by delta carries the mean shear stress
    G delta / L0 (F_t = G A delta / L0, tau = F_t / A).
The break decision must flip between 1 % below and 1 % above each of these
(md_calculateforcesdem3.cu:20-36), in the slot oracle and in the FP64
edge-list oracle alike."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.oracle import OracleLayout, lib, ptr
from rfd.params import Material
from rfd.reflayout import Grid
from test_oracle_pins import Pair, mat

BR = 0.25  # bond radius of mat()


def _broke_at(p, m, wi, wj, vj, steps, sc, tc):
    """Hold both particles in place (no integration) with the given
    velocities and evaluate the bond up to `steps` times; the evaluation at
    which it broke (0-based), or None."""
    p.set_w(0, wi)
    p.set_w(1, wj)
    p.set_v(1, vj)
    for k in range(steps):
        p.bonds(m, sc=sc, tc=tc)
        if p.ILtype[0] == 0:
            return k
    return None


def _drive(p, m, wi, wj, vj, steps, sc, tc):
    """True when the bond broke on the last of `steps` evaluations and not
    before."""
    return _broke_at(p, m, wi, wj, vj, steps, sc, tc) == steps - 1


# ------------------------------------------------- slot oracle (reference)
@pytest.mark.parametrize("side", [-1, 1])
def test_pure_bend_breaks_at_beam_stress(side):
    """Ends turned in opposite senses about y (W_i + W_j = 0: no shear, no
    frame turn): relative rotation theta = 2 w dt k; surface stress
    E theta b / L0 (tensile criterion Sc, md_calculateforcesdem3.cu:24)."""
    m = mat()
    w, k = 2e-3, 12
    theta = 2 * w * m.dt * k
    sigma = m.E * theta * BR / 1.0
    sc = sigma * (1 + 0.01 * side)
    p = Pair((1.0, 0, 0), L0=1.0)
    broke = _drive(p, m, (0, -w, 0), (0, w, 0), (0, 0, 0), k, sc, 1e9)
    assert broke == (side < 0)
    if side > 0:  # the stored bending moment is E theta / (4 L0)
        assert p.st["Mijt"][1] == pytest.approx(m.E * theta / 4.0, rel=1e-4)
        assert np.all(p.st["Oijt"] == 0)


@pytest.mark.parametrize("side", [-1, 1])
def test_pure_twist_breaks_at_torsion_stress(side):
    """Ends turned in opposite senses about the bond axis x: twist
    theta = 2 w dt k, surface shear stress G theta b / L0 (shear criterion
    Mc, md_calculateforcesdem3.cu:25)."""
    m = mat()
    w, k = 2e-3, 12
    theta = 2 * w * m.dt * k
    tau = m.G * theta * BR / 1.0
    tc = tau * (1 + 0.01 * side)
    p = Pair((1.0, 0, 0), L0=1.0)
    broke = _drive(p, m, (-w, 0, 0), (w, 0, 0), (0, 0, 0), k, 1e9, tc)
    assert broke == (side < 0)
    if side > 0:
        assert p.st["Mijn"][0] == pytest.approx(m.G * theta / 2.0, rel=1e-4)


@pytest.mark.parametrize("side", [-1, 1])
def test_pure_shear_breaks_at_mean_shear_stress(side):
    """End j slides along y by delta = v dt k: shear stress G delta / L0."""
    m = mat()
    v, k = 1e-3, 15
    delta = v * m.dt * k
    tau = m.G * delta / 1.0
    tc = tau * (1 + 0.01 * side)
    p = Pair((1.0, 0, 0), L0=1.0)
    broke = _drive(p, m, (0, 0, 0), (0, 0, 0), (0, v, 0), k, 1e9, tc)
    assert broke == (side < 0)


def test_bend_and_stretch_add_in_the_tensile_criterion():
    """Axial strain eps and a bend theta together: sigma = E eps + E theta b
    / L0; each alone stays 37.5 % below the threshold, together they break."""
    m = mat()
    w, k = 2e-3, 12
    theta = 2 * w * m.dt * k
    s_bend = m.E * theta * BR
    eps = s_bend / m.E  # same stress from stretch as from bending
    sc = 1.6 * s_bend
    p = Pair((1.0 + eps, 0, 0), L0=1.0)
    assert _broke_at(p, m, (0, -w, 0), (0, w, 0), (0, 0, 0), k, sc, 1e9) is not None
    q = Pair((1.0, 0, 0), L0=1.0)
    assert _broke_at(q, m, (0, -w, 0), (0, w, 0), (0, 0, 0), k, sc, 1e9) is None
    q2 = Pair((1.0 + eps, 0, 0), L0=1.0)
    assert _broke_at(q2, m, (0, 0, 0), (0, 0, 0), (0, 0, 0), k, sc, 1e9) is None


# --------------------------------------------- FP64 edge-list oracle (both modes)
class NatPair:
    """One bond 0-1 along x in the FP64 edge-list oracle."""

    def __init__(self, L=1.0, L0=1.0):
        self.R = np.array([0.0, L, 0.0, 0.0, 0.0, 0.0])  # SoA x0 x1 y0 y1 z0 z1
        self.V = np.zeros(6)
        self.W = np.zeros(6)
        self.st = np.zeros(8)
        self.act = np.ones(1, np.uint8)
        self.iL = np.array([1.0 / L0])
        self.F = np.zeros(6)
        self.M = np.zeros(6)

    def run(self, m, wi, wj, vj, steps, sc, tc, refmode):
        for d in range(3):
            self.W[2 * d], self.W[2 * d + 1] = wi[d], wj[d]
            self.V[2 * d + 1] = vj[d]
        bi, bj = np.array([0], np.uint32), np.array([1], np.uint32)
        br, s, t = np.array([BR]), np.array([sc]), np.array([tc])
        for k in range(steps):
            lib().orc_native_bonds_f64(ptr(self.R), ptr(self.V), ptr(self.W), 2, ptr(bi),
                                       ptr(bj), ptr(self.iL), ptr(br), ptr(s), ptr(t),
                                       ptr(self.st), ptr(self.act), 1, m.dt, m.E, m.G,
                                       refmode, ptr(self.F), ptr(self.M))
            # sliding end: positions advance by V dt between evaluations, as
            # the edge-list state assumes
            self.R = self.R + self.V * m.dt
            if not self.act[0]:
                return k == steps - 1
        return False


@pytest.mark.parametrize("refmode", [0, 1])
@pytest.mark.parametrize("side", [-1, 1])
def test_native_bend_twist_shear_thresholds(refmode, side):
    m = mat()
    w, k = 2e-3, 12
    theta = 2 * w * m.dt * k
    f = 1 + 0.01 * side
    # bend about y: tensile criterion at E theta b / L0
    p = NatPair()
    assert p.run(m, (0, -w, 0), (0, w, 0), (0, 0, 0), k, m.E * theta * BR * f, 1e9,
                 refmode) == (side < 0)
    # twist about x: shear criterion at G theta b / L0
    q = NatPair()
    assert q.run(m, (-w, 0, 0), (w, 0, 0), (0, 0, 0), k, 1e9, m.G * theta * BR * f,
                 refmode) == (side < 0)
    if side > 0:
        assert q.st[3] == pytest.approx(m.G * theta / 2.0, rel=1e-9)
    # slide along y: shear criterion at G delta / L0 (delta as the end moves)
    v, k2 = 1e-3, 15
    s = NatPair()
    assert s.run(m, (0, 0, 0), (0, 0, 0), (0, v, 0), k2, 1e9, m.G * v * m.dt * k2 * f,
                 refmode) == (side < 0)


def test_native_bending_moment_factor_by_mode():
    """The applied bending moment: J Mt (J = pi b^4) in the physical mode,
    times the extra dt E / (4 L0) of md_calculateforcesdem3.cu:148 in the
    reference mode; the stored Mt is E theta / (4 L0) in both."""
    m = mat()
    w, k = 2e-3, 12
    theta = 2 * w * m.dt * k
    J = math.pi * BR ** 4
    for refmode, fac in ((0, 1.0), (1, m.dt * 0.25 * m.E)):
        p = NatPair()
        p.run(m, (0, -w, 0), (0, w, 0), (0, 0, 0), k, 1e9, 1e9, refmode)
        assert p.st[5] == pytest.approx(m.E * theta / 4.0, rel=1e-9)
        # particle 0's moment about y (SoA index 2 * 1 + 0)
        assert p.M[2] == pytest.approx(J * fac * m.E * theta / 4.0, rel=1e-9)
        assert p.M[3] == pytest.approx(-p.M[2], rel=1e-12)


# ------------------------------------------------------- reference FIRE step
def test_fire_step_discards_the_old_velocity():
    """d_CalculateIncrementsFIRE (md_firerelaxation.cu:28-30) starts from
    v = F dt/m, not from the stored V: the new V and R do not depend on the
    V passed in.  v is parallel to F, so the mixing (1 - a) v + a |v| F/|F|
    returns v itself: V = F dt/m up to rounding, R += V dt."""
    g = np.random.default_rng(4)
    N = 40
    F = g.standard_normal(3 * N).astype(np.float32)
    R0 = g.uniform(0, 5, 3 * N).astype(np.float32)
    dt_m, dt, alpha = np.float32(0.03), np.float32(0.2), np.float32(0.1)
    outs = []
    for vin in (np.zeros(3 * N), g.standard_normal(3 * N), 1e3 * g.standard_normal(3 * N)):
        V = vin.astype(np.float32)
        R = R0.copy()
        lib().orc_fire_increments(ptr(F), ptr(V), ptr(R), N, dt_m, dt, alpha)
        outs.append((V, R))
    for V, R in outs[1:]:
        assert np.array_equal(V, outs[0][0]) and np.array_equal(R, outs[0][1])
    V, R = outs[0]
    np.testing.assert_allclose(V, F.astype(np.float64) * float(dt_m), rtol=1e-6)
    np.testing.assert_allclose(R, R0 + V.astype(np.float64) * float(dt), rtol=1e-6, atol=1e-7)
    # a particle without load stops, whatever it carried
    Fz = np.zeros(3, np.float32)
    Vz = np.array([5.0, -3.0, 2.0], np.float32)
    Rz = np.ones(3, np.float32)
    lib().orc_fire_increments(ptr(Fz), ptr(Vz), ptr(Rz), 1, dt_m, dt, alpha)
    assert np.all(Vz == 0) and np.all(Rz == 1)


def test_fire_step_small_force_guard():
    """|F|^2 <= 1e-12 (md_firerelaxation.cu:35-42): the mixing weight becomes
    alpha |v| 1e6 instead of alpha |v| / |F|.  F = 1e-7 along x, dt/m = 0.03,
    alpha = 0.1: v = 3e-9, V = 0.9 * 3e-9 + 0.1 * 3e-9 * 1e6 * 1e-7
    = 2.73e-9 (by hand)."""
    F = np.array([1e-7, 0.0, 0.0], np.float32)
    V = np.array([1.0, 1.0, 1.0], np.float32)
    R = np.zeros(3, np.float32)
    lib().orc_fire_increments(ptr(F), ptr(V), ptr(R), 1, np.float32(0.03), np.float32(2.0),
                              np.float32(0.1))
    assert V[0] == pytest.approx(2.73e-9, rel=1e-5)
    assert V[1] == 0 and V[2] == 0
    assert R[0] == pytest.approx(5.46e-9, rel=1e-5)


# ----------------------------------------------------------- 8-cell slot order
# particle 3 sits in cell (1,1,1) of a 4x4x4 grid of edge 2.9, near its +x,
# -y and +z faces; one partner in each of the 8 cells the reference searches,
# with ids that are not in cell order
_I = 3
_PARTNERS = {  # cell step (dx, dy, dz) -> (id, position)
    (0, 0, 0): (8, (5.0, 3.5, 5.0)),
    (1, 0, 0): (1, (6.0, 3.2, 5.5)),
    (0, -1, 0): (6, (5.5, 2.7, 5.5)),
    (0, 0, 1): (2, (5.5, 3.2, 6.0)),
    (1, -1, 0): (7, (6.0, 2.7, 5.5)),
    (1, 0, 1): (0, (6.0, 3.2, 6.0)),
    (0, -1, 1): (5, (5.5, 2.7, 6.0)),
    (1, -1, 1): (4, (6.0, 2.7, 6.0)),
}
# md_interaction_list_construct.cu:60-83: own, x, y, z, xy, xz, yz, xyz
_ORDER = [(0, 0, 0), (1, 0, 0), (0, -1, 0), (0, 0, 1), (1, -1, 0), (1, 0, 1), (0, -1, 1),
          (1, -1, 1)]


def eight_cell_case():
    """Positions (SoA, float32), grid, cut radius^2 and the expected row of
    particle 3 for the 8-cell order test."""
    a = np.float32(2.9)
    grid = Grid(float(a), float(np.float32(1) / a), 4, 4, 4)
    P = np.zeros((9, 3), np.float32)
    P[_I] = (5.5, 3.2, 5.5)
    for step, (pid, pos) in _PARTNERS.items():
        P[pid] = pos
        ci = np.floor(np.array(pos, np.float32) * np.float32(grid.inv_a)).astype(int)
        assert tuple(ci - np.array([1, 1, 1])) == step
    aacut = float(np.float32(1.45) ** 2)
    want = [_PARTNERS[s][0] for s in _ORDER]
    return np.ascontiguousarray(P.T).reshape(-1), grid, aacut, want


def test_il_construct_full_eight_cell_order():
    R, grid, aacut, want = eight_cell_case()
    o = OracleLayout(R, 10, grid)
    o.cell_distribution()
    o.il_construct(aacut)
    assert o.last_status == 0
    assert o.IL[_I * 10:_I * 10 + 10].tolist() == want + [0xFFFFFFFF] * 2


def test_il_reconstruct_full_eight_cell_order():
    """The reconstruct inserts new partners in the same visiting order
    (md_interaction_list_reconstructbpm.cu:80-139), each at the first free
    slot; a partner already listed keeps its slot."""
    R, grid, aacut, want = eight_cell_case()
    o = OracleLayout(R, 10, grid)
    o.cell_distribution()
    row = slice(_I * 10, _I * 10 + 10)
    o.IL[row] = [0xFFFFFFFF, 0xFFFFFFFF, want[3], 0xFFFFFFFF] + [0xFFFFFFFF] * 6
    o.il_reconstruct_bpm(aacut, quirk_p0=True)
    assert o.last_status == 0
    rest = [x for x in want if x != want[3]]
    assert o.IL[row].tolist() == rest[:2] + [want[3]] + rest[2:] + [0xFFFFFFFF] * 2


# ------------------------------------------------------ parameter derivation
def test_granite_parameters_by_hand():
    """potential_data::CalculateParameters (md_data_types.h:153-185) for the
    calibration start (mp_determineconjgradient.cu:15-33), against values
    worked out by hand in the textbook forms: G = E / (2 (1 + nu));
    m = rho 4/3 pi r^3; I = 2/5 m r^2 (solid sphere); Hertz E* = E / (2 (1 -
    nu^2)) and Mindlin G* = G / (2 (2 - nu)) for two equal spheres; Rayleigh
    step pi r sqrt(rho / G) / (0.163 nu + 0.8766); the spring step 2 / omega
    with omega = sqrt(k / m) for the stiffer bond spring k = E A / (2 r);
    dt one tenth of the smaller."""
    g = Material.reference_granite()
    assert g.E == pytest.approx(195.255)
    assert g.rho == pytest.approx(2.61)
    assert g.p_r == pytest.approx(0.5)
    assert g.b_r == pytest.approx(0.1650075)
    assert g.b_d == pytest.approx(1.0593169)
    assert g.G == pytest.approx(78.28989575, rel=1e-9)
    assert g.p_m == pytest.approx(1.366592804, rel=1e-9)
    assert g.p_I == pytest.approx(0.1366592804, rel=1e-9)
    assert g.hm_E == pytest.approx(103.9706451, rel=1e-8)
    assert g.hm_G == pytest.approx(22.33026120, rel=1e-8)
    assert g.b_S == pytest.approx(0.08553763561, rel=1e-9)
    assert g.T_rayleigh == pytest.approx(0.3128122655, rel=1e-9)
    assert g.T_sim == pytest.approx(0.5720972468, rel=1e-9)
    assert g.dt == pytest.approx(0.03128122655, rel=1e-9)
    assert g.dt_d_m == pytest.approx(0.03128122655 / 1.366592804, rel=1e-9)
    assert g.dt_d_I == pytest.approx(0.03128122655 / 0.1366592804, rel=1e-9)
    # the restitution constant for e = 0.6: ln e / sqrt(pi^2 + ln^2 e)
    assert g.p_A == pytest.approx(-0.1604930467, rel=1e-9)


def test_time_step_rule_picks_the_smaller_period():
    """A soft, dense particle with stiff bonds: the spring period becomes the
    smaller one and sets dt (the other branch of md_data_types.h:172-175);
    the bond spring uses the larger of E and G (E here)."""
    m = Material(E=30.0, nu=0.25, rho=2.65, p_r=0.5, b_r=0.5, b_d=1.05, sc=1, tc=1)
    # k = E pi b^2 / (2 r) = 30 pi 0.25 = 23.5619; m = 2.65 pi/6 = 1.387537
    assert m.T_sim == pytest.approx(2.0 * math.sqrt(1.387537 / 23.56194), rel=1e-5)
    # Rayleigh: pi 0.5 sqrt(2.65 / 12) / (0.163 0.25 + 0.8766) = 0.73816 / 0.91735
    assert m.T_rayleigh == pytest.approx(0.7382 / 0.91735, rel=1e-3)
    assert m.T_sim < m.T_rayleigh
    assert m.dt == pytest.approx(0.1 * m.T_sim, rel=1e-12)


# ------------------------------------------------- slot contact law (FP32)
# Two equal spheres r, mass m, overlap delta = 2 r - L: Hertz normal force
# 4/3 E* sqrt(R*) delta^1.5 with R* = r / 2; Mindlin tangential stiffness
# 8 G* sqrt(R* delta); viscous damping 2 sqrt(5/6) beta sqrt(k m*) = 1.8257
# beta sqrt(k m*) with m* = m / 2 (md_calculateforcesdem3.cu:235-314).
def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


def _contact_pair(m, n, delta, O=None, wi=None):
    L = 2 * m.p_r - delta
    p = Pair(tuple(np.float32(L * n)), bonded=False)
    if O is not None:
        p.st["Oijt"][0:3] = np.asarray(O, np.float32)
    if wi is not None:  # i spins, j counter-spins: no sliding from the spins
        p.set_w(0, wi)
        p.set_w(1, -np.asarray(wi))
    return p


def test_slot_contact_spring_force_and_lever_in_general_direction():
    m = mat()
    n = _unit((0.3, -0.5, 0.8))
    t = _unit(np.cross(n, (1.0, 0.0, 0.0)))
    delta, a = 0.01, 2e-5
    wi = 3e-3 * _unit((0.2, 0.9, -0.4))
    p = _contact_pair(m, n, delta, O=a * t, wi=wi)
    p.contact(m)
    F0, M0 = p.vec(p.F, 0), p.vec(p.M, 0)
    Rs = m.p_r / 2
    fn = 4.0 / 3.0 * m.hm_E * math.sqrt(Rs) * delta ** 1.5
    Ft = 8 * m.hm_G * math.sqrt(Rs * delta) * a * t
    assert np.linalg.norm(Ft) < m.p_mu * fn  # below the Coulomb cap
    np.testing.assert_allclose(F0, -fn * n + Ft, rtol=2e-4, atol=2e-4 * fn)
    expect = m.p_r * np.cross(n, Ft) - m.p_mur * fn * wi / np.linalg.norm(wi)
    np.testing.assert_allclose(M0, expect, rtol=2e-4, atol=2e-4 * np.abs(expect).max())
    np.testing.assert_allclose(p.st["Oijt"][0:3], a * t, rtol=1e-5, atol=1e-6 * a)


def test_slot_contact_old_spring_turned_into_tangent_plane():
    """A spring with a part along n (the pair has turned since it was
    stored) is turned into the tangent plane with its length kept."""
    m = mat()
    n = _unit((0.6, 0.0, 0.8))
    t = _unit(np.cross(n, (0.0, 1.0, 0.0)))
    a, b = 2e-5, 1.5e-5
    p = _contact_pair(m, n, 0.01, O=a * t + b * n)
    p.contact(m)
    np.testing.assert_allclose(p.st["Oijt"][0:3], math.hypot(a, b) * t, rtol=1e-5,
                               atol=1e-6 * a)
    Rs = m.p_r / 2
    Kt = 8 * m.hm_G * math.sqrt(Rs * 0.01)
    F0 = p.vec(p.F, 0)
    np.testing.assert_allclose(F0 - n * (F0 @ n), Kt * math.hypot(a, b) * t, rtol=2e-4,
                               atol=1e-6)


def test_slot_contact_tangential_damping():
    """j slides along t at v from a fresh spring: F_t = Kt v dt -
    1.8257 beta sqrt(Kt m / 2) v on i."""
    m = mat()
    n = _unit((1.0, 1.0, 0.0))
    t = _unit((-1.0, 1.0, 0.0))
    v, delta = 2e-4, 0.01
    p = _contact_pair(m, n, delta)
    p.set_v(1, v * t)
    p.contact(m)
    Rs = m.p_r / 2
    Kt = 8 * m.hm_G * math.sqrt(Rs * delta)
    ft = Kt * v * m.dt - 1.8257 * m.p_A * math.sqrt(Kt * m.p_m / 2) * v
    fn = 4.0 / 3.0 * m.hm_E * math.sqrt(Rs) * delta ** 1.5
    assert abs(ft) < m.p_mu * fn
    F0 = p.vec(p.F, 0)
    assert F0 @ t == pytest.approx(ft, rel=2e-4)
    np.testing.assert_allclose(p.st["Oijt"][0:3], v * m.dt * t, rtol=1e-5, atol=1e-6 * v * m.dt)


# ------------------------------------------- native contact: lever arm of j
def test_native_contact_slip_uses_each_radius():
    """Contact point velocity V_j - V_i + n x (W_i r_i + W_j r_j): only j
    spins, so after one step from a fresh spring the spring holds
    (n x W_j) r_j dt (r_j = 0.6, not r_i = 0.4), in a general direction."""
    from oracle.oracle import native_contacts

    n = _unit((0.6, 0.0, 0.8))
    wj = 0.2 * _unit((0.2, 0.9, -0.4))
    R = np.array([[0.0, 0, 0], list(0.95 * n)])
    V = np.zeros((2, 3))
    W = np.array([[0.0, 0, 0], list(wj)])
    O = np.zeros((1, 3))
    native_contacts(R, V, W, [0.4, 0.6], np.array([[0, 1]]), [1.0], O, 1e-3, 16.0, 4.0,
                    -0.16, 0.45, 0.05, 11.1)
    np.testing.assert_allclose(O[0], np.cross(n, wj) * 0.6 * 1e-3, rtol=1e-12, atol=1e-18)


# ------------------------------------------------------------ platens, prune
def test_platen_uniaxial_clamp_velocity_and_sign_filter():
    """md_uniaxialcompression3_simple.cu:19-76: a top-band particle pulled
    down (F_z < 0) adds nothing to the top sum; one past Zt takes the
    platen's velocity vt; a bottom-band one pushed up adds nothing."""
    P = np.array([[0, 0, 9.8], [0, 0, 10.3], [0, 0, 0.2], [0, 0, -0.2]], np.float32)
    N = len(P)
    R = np.ascontiguousarray(P.T).reshape(-1).copy()
    V = np.zeros(3 * N, np.float32)
    F = np.zeros(3 * N, np.float32)
    F[2 * N:] = [-2.0, 3.0, 4.0, -1.0]
    sums = (C.c_double * 2)()
    vt = np.float32(-2.5e-3)
    lib().orc_platen_uniaxial(ptr(R), ptr(V), ptr(F), N, 0.0, 0.0, 4.0, 10.0, 0.0, 9.5, 0.5,
                              vt, 100.0, sums)
    assert (sums[0], sums[1]) == (-1.0, 3.0)
    assert R[2 * N + 1] == 10.0 and V[2 * N + 1] == vt and F[2 * N + 1] == 0
    assert R[2 * N + 3] == 0.0 and V[2 * N + 3] == 0 and F[2 * N + 3] == 0
    assert F[2 * N + 0] == -2.0 and F[2 * N + 2] == 4.0


def test_delete_far_clears_every_slot_array():
    """d_DeleteFarLinks (md_interaction_list_reconstructbpm.cu:141-186):
    a pruned slot gets the sentinel, type 0 and zero rij, Oijt, F, M terms;
    1/L0 stays (the reference clears a local pointer, :173)."""
    from oracle.oracle import OracleLayout

    R = np.array([0.0, 5.0, 0.0, 0.0, 0.0, 0.0], np.float32)  # pair 5 apart along x
    o = OracleLayout(R, 1, Grid(1, 1, 1, 1, 1))
    o.IL[:] = [1, 0]
    o.ILtype[:] = 1
    o.iL[:] = 0.25
    for k in o.slot:
        o.slot[k][:] = 0.5
    o.delete_far_links(4.0)  # (far cut 2)^2 < 25
    assert o.IL.tolist() == [0xFFFFFFFF] * 2 and o.ILtype.tolist() == [0, 0]
    for k in o.slot:
        assert np.all(o.slot[k] == 0), k
    assert np.all(o.iL == 0.25)


def test_fire_power_with_rotations_known_answer():
    """d_FdotVEntire (md_firerelaxation.cu:372) sums F.V; with rotations the
    build adds M.W, |M|^2 and |W|^2 (hand sums on two particles)."""
    N = 2
    F = np.array([1, 2, 3, 4, 5, 6], np.float32)  # SoA x0 x1 y0 y1 z0 z1
    V = np.array([1, 0, -1, 0, 2, 1], np.float32)
    M = np.array([0, 1, 0, 0, 2, 0], np.float32)
    W = np.array([3, 1, 0, 0, 1, 4], np.float32)
    out = (C.c_double * 4)()
    lib().orc_fire_power(ptr(F), ptr(V), ptr(M), ptr(W), N, out)
    # F.V = 1 - 3 + 10 + 6 = 14; M.W = 1 + 2 = 3
    assert out[0] == 17.0
    # |F|^2 = 91, |M|^2 = 5; |V|^2 = 7, |W|^2 = 27; max_i |F_i|^2 = 4+16+36 = 56
    assert out[1] == 96.0 and out[2] == 34.0 and out[3] == 56.0
    lib().orc_fire_power(ptr(F), ptr(V), None, None, N, out)
    assert (out[0], out[1], out[2], out[3]) == (14.0, 91.0, 7.0, 56.0)
