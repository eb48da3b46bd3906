"""Pins for the CPU oracle: closed forms, conservation laws, brute force and
special cases chosen so that a dropped term, a wrong sign or a swapped index
in the restatement fails at least one of them.  (The reference ships no test
vectors or fixtures; these are what the mathematics fixes.)"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle.oracle import lib, ptr
from rfd import cases
from rfd.params import Compression, Material

PI_F = np.float32(3.14159265358979)


def mat():
    return Material(E=30.0, nu=0.25, rho=2.65, p_r=0.5, b_r=0.25, b_d=1.05, sc=1e9, tc=1e9)


class Pair:
    """Two particles i=0, j=1 with one slot each pointing at the other."""

    def __init__(self, r, bonded=True, L0=None, K=1):
        self.N, self.K = 2, K
        self.R = np.zeros(6, np.float32)
        self.R[1], self.R[3], self.R[5] = r  # particle 1 at r, particle 0 at origin
        self.V = np.zeros(6, np.float32)
        self.W = np.zeros(6, np.float32)
        self.F = np.zeros(6, np.float32)
        self.M = np.zeros(6, np.float32)
        S = 2 * K
        self.IL = np.full(S, 0xFFFFFFFF, np.uint32)
        self.IL[0], self.IL[K] = 1, 0
        self.ILtype = np.zeros(S, np.uint8)
        if bonded:
            self.ILtype[0] = self.ILtype[K] = 1
        L0 = L0 if L0 is not None else float(np.linalg.norm(np.asarray(r, np.float64)))
        self.iL = np.full(S, np.float32(1.0 / L0), np.float32)
        self.st = {k: np.zeros(3 * S, np.float32) for k in ("rij", "Oijt", "Mijn", "Mijt")}

    def set_v(self, p, v):
        for d in range(3):
            self.V[p + 2 * d] = v[d]

    def set_w(self, p, w):
        for d in range(3):
            self.W[p + 2 * d] = w[d]

    def vec(self, a, p):
        return np.array([a[p], a[p + 2], a[p + 4]], np.float64)

    def bonds(self, m, sc=None, tc=None):
        self.F[:] = 0
        self.M[:] = 0
        lib().orc_forces_bpm(ptr(self.R), ptr(self.V), ptr(self.W), ptr(self.IL),
                             ptr(self.ILtype), ptr(self.iL), ptr(self.st["rij"]),
                             ptr(self.st["Oijt"]), ptr(self.st["Mijn"]), ptr(self.st["Mijt"]),
                             ptr(self.F), ptr(self.M), 2, self.K, m.dt, m.E, m.G, m.b_r,
                             m.sc if sc is None else sc, m.tc if tc is None else tc)

    def contact(self, m):
        self.F[:] = 0
        self.M[:] = 0
        lib().orc_forces_contact(ptr(self.R), ptr(self.V), ptr(self.W), ptr(self.IL),
                                 ptr(self.ILtype), ptr(self.st["Oijt"]), ptr(self.F),
                                 ptr(self.M), 2, self.K, m.dt, m.hm_E, m.hm_G, m.p_A, m.p_mu,
                                 m.p_mur, m.p_m, m.p_r)


# ----------------------------------------------------------------- helpers
def test_norm3_and_rsqrt_correctly_rounded():
    g = np.random.default_rng(0)
    x = g.standard_normal((2000, 3)).astype(np.float32) * np.float32(3.7)
    half_ulp_ok = 0
    for a, b, c in x:
        f = lib().orc_norm3(a, b, c)
        S = sum(Fraction(float(v)) ** 2 for v in (a, b, c))
        ulp = np.spacing(np.float32(f))
        # |f - sqrt(S)| <= 1 ulp, checked with exact rational arithmetic
        lo, hi = Fraction(float(f)) - Fraction(float(ulp)), Fraction(float(f)) + Fraction(float(ulp))
        assert lo * lo <= S <= hi * hi
        lo2 = Fraction(float(f)) - Fraction(float(ulp)) / 2
        hi2 = Fraction(float(f)) + Fraction(float(ulp)) / 2
        half_ulp_ok += lo2 * lo2 <= S <= hi2 * hi2
    assert half_ulp_ok == len(x)
    for v in g.uniform(1e-6, 1e6, 2000).astype(np.float32):
        r = lib().orc_rsqrt(v)
        assert abs(r - 1 / math.sqrt(float(v))) <= 0.5 * np.spacing(np.float32(r)) * 1.0001


# ---------------------------------------------------------------- bond law
def test_axial_stretch_closed_form():
    m = mat()
    p = Pair((1.001, 0, 0), L0=1.0)
    p.bonds(m)
    F0, F1 = p.vec(p.F, 0), p.vec(p.F, 1)
    area = float(PI_F) * m.b_r ** 2
    want = m.E * area * 0.001
    assert F0[0] == pytest.approx(want, rel=2e-4)  # eps=1e-3 in float32
    assert abs(F0[1]) + abs(F0[2]) == 0
    assert np.array_equal(F1, -F0)
    np.testing.assert_array_equal(p.st["rij"][:3], [1, 0, 0])


def test_rest_length_gives_zero_force():
    m = mat()
    p = Pair((0.0, 1.0, 0.0), L0=1.0)
    p.bonds(m)
    assert np.all(p.F == 0) and np.all(p.M == 0)


def test_shear_displacement_closed_form():
    """j slides along +y at v for k steps: |Ft| = G*pi*b^2*delta/L0 on i, +y."""
    m = mat()
    p = Pair((1.0, 0, 0), L0=1.0)
    v, k = 1e-3, 20
    p.set_v(1, (0, v, 0))
    for _ in range(k):
        p.bonds(m)
    delta = v * m.dt * k
    want = m.G * float(PI_F) * m.b_r ** 2 * delta / 1.0
    F0 = p.vec(p.F, 0)
    assert F0[1] == pytest.approx(want, rel=1e-4)
    # 0.5 r x Ft moment on both ends, about +z
    M0 = p.vec(p.M, 0)
    assert M0[2] == pytest.approx(0.5 * 1.0 * want, rel=1e-4)


def test_twist_closed_form():
    """j spins about the bond axis: stored Mn = G*theta/(2 L0), moment on i =
    G*J*theta/L0 with J = pi b^4 / 2; j gets the opposite moment."""
    m = mat()
    p = Pair((1.0, 0, 0), L0=1.0)
    w, k = 2e-3, 25
    p.set_w(1, (w, 0, 0))
    for _ in range(k):
        p.bonds(m)
    theta = w * m.dt * k
    assert p.st["Mijn"][0] == pytest.approx(m.G * theta / 2.0, rel=1e-4)
    J = float(PI_F) * m.b_r ** 4 / 2.0
    M0, M1 = p.vec(p.M, 0), p.vec(p.M, 1)
    assert M0[0] == pytest.approx(m.G * J * theta / 1.0, rel=1e-4)
    assert M1[0] == pytest.approx(-M0[0], rel=1e-6)
    assert np.all(p.F == 0)


def test_bend_closed_form_with_reference_factor():
    """j rotates about +y: stored Mt = E*theta/(4 L0); the applied moment
    carries the extra dt*E/(4 L0) factor of md_calculateforcesdem3.cu:148;
    the rotation also shears the bond (O along +z) which adds 0.5 r x Ft."""
    m = mat()
    p = Pair((1.0, 0, 0), L0=1.0)
    w, k = 2e-3, 10
    p.set_w(1, (0, w, 0))
    for _ in range(k):
        p.bonds(m)
    theta = w * m.dt * k
    assert p.st["Mijt"][1] == pytest.approx(m.E * theta / 4.0, rel=1e-4)
    area = float(PI_F) * m.b_r ** 2
    J4 = float(PI_F) * m.b_r ** 4
    O = 0.5 * w * 1.0 * m.dt * k       # (w/2 y) x (L x) = -(wL/2) z, minus sign -> +z
    Ft = m.G * area * O
    F0 = p.vec(p.F, 0)
    assert F0[2] == pytest.approx(Ft, rel=1e-4)
    Mt_applied = (m.E * theta / 4.0) * (m.dt * 0.25 * m.E) * J4
    M0 = p.vec(p.M, 0)
    assert M0[1] == pytest.approx(Mt_applied - 0.5 * 1.0 * Ft, rel=1e-4)


def test_newton3_and_angular_momentum():
    g = np.random.default_rng(3)
    m = mat()
    for trial in range(20):
        r = g.uniform(-1, 1, 3)
        r = r / np.linalg.norm(r) * g.uniform(0.9, 1.1)
        p = Pair(tuple(r), L0=1.0)
        p.set_v(0, g.standard_normal(3) * 1e-3)
        p.set_v(1, g.standard_normal(3) * 1e-3)
        p.set_w(0, g.standard_normal(3) * 1e-3)
        p.set_w(1, g.standard_normal(3) * 1e-3)
        for _ in range(3):
            p.bonds(m)
        F0, F1 = p.vec(p.F, 0), p.vec(p.F, 1)
        assert np.array_equal(F0, -F1), trial
        R0, R1 = p.vec(p.R, 0), p.vec(p.R, 1)
        L = np.cross(R0, F0) + np.cross(R1, F1) + p.vec(p.M, 0) + p.vec(p.M, 1)
        scale = np.abs(F0).max() + np.abs(p.vec(p.M, 0)).max()
        assert np.abs(L).max() <= 1e-5 * scale, trial


def test_rigid_body_motion_is_force_free():
    """Rigid translation + rotation: no relative velocity, no bond load."""
    m = mat()
    p = Pair((0.6, 0.8, 0.0), L0=1.0)
    Om = np.array([1e-3, -2e-3, 3e-3])
    V0 = np.array([5e-3, 1e-3, -2e-3])
    for q in (0, 1):
        Rq = p.vec(p.R, q)
        p.set_v(q, V0 + np.cross(Om, Rq))
        p.set_w(q, Om)
    p.bonds(m)
    assert np.abs(p.F).max() < 1e-9 and np.abs(p.M).max() < 1e-9


@pytest.mark.parametrize("below", [True, False])
def test_break_threshold(below):
    m = mat()
    p = Pair((1.002, 0, 0), L0=1.0)
    Sc = np.float32(m.E) * (np.float32(1.002) - np.float32(1.0)) * np.float32(1.0)
    sc = float(np.nextafter(Sc, np.float32(0))) if below else float(np.nextafter(Sc, np.float32(1)))
    p.st["Oijt"][:3] = 1.0  # state that must be wiped on a break
    p.bonds(m, sc=sc)
    broken = p.ILtype[0] == 0
    assert broken == below
    if below:
        assert np.all(p.st["Oijt"][:3] == 0) and np.all(p.st["Mijn"][:3] == 0)
        # the force of the breaking step is still applied (md_calculateforcesdem3.cu:450,472)
        assert p.F[0] > 0


# ------------------------------------------------------------- contact law
def test_hertz_normal_closed_form():
    m = mat()
    delta = 0.01
    p = Pair((2 * m.p_r - delta, 0, 0), bonded=False)
    p.contact(m)
    F0 = p.vec(p.F, 0)
    want = -(4.0 / 3.0) * m.hm_E * math.sqrt(m.p_r / 2.0) * delta ** 1.5
    assert F0[0] == pytest.approx(want, rel=1e-4)
    assert np.array_equal(p.vec(p.F, 1), -F0)


def test_no_contact_beyond_touching():
    m = mat()
    p = Pair((2 * m.p_r + 1e-3, 0, 0), bonded=False)
    p.st["Oijt"][:] = 1.0
    p.contact(m)
    assert np.all(p.F == 0) and np.all(p.st["Oijt"] == 0)


def test_coulomb_cap():
    m = mat()
    delta = 0.01
    p = Pair((2 * m.p_r - delta, 0, 0), bonded=False)
    p.st["Oijt"][0:3] = [0, 0.05, 0]  # large tangential spring of slot 0->1
    p.set_w(0, (0, 0, 1e-3))          # so that the rolling branch does not wipe ms
    p.contact(m)
    F0 = p.vec(p.F, 0)
    fn = abs(F0[0])
    ft = math.hypot(F0[1], F0[2])
    assert ft == pytest.approx(m.p_mu * fn, rel=1e-5)


def test_rolling_torque_and_wipe_quirk():
    m = mat()
    delta = 0.01
    p = Pair((2 * m.p_r - delta, 0, 0), bonded=False)
    w = np.array([0.0, 0.0, 3e-3])
    p.set_w(0, w)
    p.contact(m)
    F0, M0 = p.vec(p.F, 0), p.vec(p.M, 0)
    n = np.array([1.0, 0, 0])
    fn = abs(F0 @ n)
    Ft = F0 - n * (F0 @ n)
    expect = m.p_r * np.cross(n, Ft) - m.p_mur * fn * w / np.linalg.norm(w)
    np.testing.assert_allclose(M0, expect, rtol=1e-4, atol=1e-12)
    # W_i == 0: the reference's else-branch wipes the whole accumulator
    q = Pair((2 * m.p_r - delta, 0, 0), bonded=False)
    w1 = np.array([0.0, 0.0, 3e-3])
    q.set_w(1, w1)
    q.contact(m)
    assert np.all(q.vec(q.M, 0) == 0)
    # particle 1 spins, so its own row keeps the lever moment of its
    # tangential force (seen along n1 = -x) and gets the rolling torque
    F1, M1 = q.vec(q.F, 1), q.vec(q.M, 1)
    n1 = np.array([-1.0, 0, 0])
    fn1 = abs(F1 @ n1)
    Ft1 = F1 - n1 * (F1 @ n1)
    assert np.linalg.norm(Ft1) > 0
    expect1 = m.p_r * np.cross(n1, Ft1) - m.p_mur * fn1 * w1 / np.linalg.norm(w1)
    np.testing.assert_allclose(M1, expect1, rtol=1e-4, atol=1e-12)


def test_head_on_restitution():
    """The damping law (-1.8257 alpha sqrt(k m/2)) gives e close to the
    material's restitution coefficient for a head-on Hertz impact."""
    m = mat()
    p = Pair((2 * m.p_r + 0.002, 0, 0), bonded=False)
    v0 = 1e-3
    p.set_v(1, (-v0, 0, 0))
    for _ in range(200000):
        p.contact(m)
        lib().orc_increments(ptr(p.F), ptr(p.V), ptr(p.R), ptr(p.M), ptr(p.W), 2,
                             m.dt_d_m, m.dt, m.dt_d_I)
        rel = p.V[1] - p.V[0]
        if rel > 0 and p.R[1] - p.R[0] > 2 * m.p_r:
            break
    e = (p.V[1] - p.V[0]) / v0
    assert e == pytest.approx(m.p_e, abs=0.06)


def _beta(e):
    # damping ratio of the restitution coefficient e (Tsuji, Tanaka and
    # Ishida 1992): ln e / sqrt(ln^2 e + pi^2), negative for e < 1
    le = math.log(e)
    return le / math.sqrt(le * le + math.pi * math.pi)


@pytest.mark.parametrize("u", [2e-2, -2e-2])
def test_normal_damping_tsuji_closed_form(u):
    """Hertz spring plus Tsuji's viscous normal force for equal spheres:
    F = -(4/3) E* sqrt(R*) d^1.5 - 2 sqrt(5/6) |beta| sqrt(S_n m*) u along the
    line of centres, S_n = 2 E* sqrt(R* d), R* = r/2, m* = m/2, u the approach
    speed (negative: separating).  The damping opposes the relative motion."""
    m = mat()
    delta = 2.0 ** -7  # 2 r - delta is exact in FP32
    p = Pair((2 * m.p_r - delta, 0, 0), bonded=False)
    p.set_v(1, (-u, 0, 0))  # j moves toward i for u > 0
    p.contact(m)
    E_star = m.E / (2.0 * (1.0 - m.nu ** 2))
    Rs = m.p_r / 2.0
    Sn = 2.0 * E_star * math.sqrt(Rs * delta)
    spring = (4.0 / 3.0) * E_star * math.sqrt(Rs) * delta ** 1.5
    damp = 2.0 * math.sqrt(5.0 / 6.0) * abs(_beta(m.p_e)) * math.sqrt(Sn * m.p_m / 2.0) * u
    assert abs(damp) / spring > 0.1  # the damping term is not lost in round-off
    F0 = p.vec(p.F, 0)
    # md_calculateforcesdem3.cu:245 writes 2 sqrt(5/6) = 1.825742 as 1.8257
    # (2.3e-5 relative); FP32 round-off adds about 1e-6 of each term
    assert F0[0] == pytest.approx(-(spring + damp), abs=5e-5 * (spring + abs(damp)))
    assert F0[1] == 0 and F0[2] == 0
    assert np.array_equal(p.vec(p.F, 1), -F0)


def _rotation(seed):
    if seed is None:
        return np.eye(3)
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


@pytest.mark.parametrize("seed", [None, 1, 2, 3])
@pytest.mark.parametrize("wj_sign,slip", [(-1.0, 0.0), (0.0, 1.0), (1.0, 2.0)])
def test_contact_point_slip_kinematics(wj_sign, slip, seed):
    """Sphere i at the origin spins at w about z; j touches it along +x with
    spin wj_sign * w.  The surface of i at the contact point moves at r w (+y),
    that of j at -wj_sign r w (+y), so counter-rotating spheres roll on each
    other without slip, and otherwise j's surface slips past i's at slip * r w
    along -y.  One call from a zero spring: the tangential force on i is
    Mindlin's spring over dt of that slip plus Tsuji's tangential damping,
    S_t = 8 G* sqrt(R* d), along -y, and its moment r n x F_t opposes the spin.
    The same holds with the whole configuration turned by a rotation Q (seeded;
    None: no turn), which puts every component of the slip velocity to use."""
    m = mat()
    delta = 2.0 ** -7
    w = 3e-3
    Q = _rotation(seed)
    p = Pair(tuple(Q @ np.array([2 * m.p_r - delta, 0.0, 0.0])), bonded=False)
    p.set_w(0, Q @ np.array([0.0, 0.0, w]))
    p.set_w(1, Q @ np.array([0.0, 0.0, wj_sign * w]))
    p.contact(m)
    G_star = (m.E / (2.0 * (1.0 + m.nu))) / (2.0 * (2.0 - m.nu))
    St = 8.0 * G_star * math.sqrt(m.p_r / 2.0 * delta)
    vs = slip * m.p_r * w
    ft = St * vs * m.dt + 2.0 * math.sqrt(5.0 / 6.0) * abs(_beta(m.p_e)) * math.sqrt(
        St * m.p_m / 2.0) * vs
    F0, M0 = Q.T @ p.vec(p.F, 0), Q.T @ p.vec(p.M, 0)  # back in the unturned frame
    # FP32 positions of a turned pair carry round-off of about 1e-7 of r
    tol = 0.0 if seed is None else 2e-4 * abs(F0[0])
    if slip == 0.0:
        assert abs(F0[1]) <= tol and abs(F0[2]) <= tol
    else:
        # md_calculateforcesdem3.cu:299 writes 2 sqrt(5/6) as 1.8257, as above
        assert F0[1] == pytest.approx(-ft, rel=5e-5 if seed is None else 2e-3)
        assert abs(F0[2]) <= tol
        # the friction moment on i works against its spin
        assert M0[2] < -m.p_r * ft * 0.99


# ------------------------------------------------------------ integration
def test_increments_exact_affine():
    g = np.random.default_rng(5)
    N = 50
    F, V, R, M, W = (g.standard_normal(3 * N).astype(np.float32) for _ in range(5))
    dt_m, dt, dt_I = np.float32(0.3), np.float32(0.01), np.float32(0.7)
    Vx = V + F * dt_m
    Rx = R + Vx * dt
    Wx = W + M * dt_I
    lib().orc_increments(ptr(F), ptr(V), ptr(R), ptr(M), ptr(W), N, dt_m, dt, dt_I)
    assert np.array_equal(V, Vx) and np.array_equal(R, Rx) and np.array_equal(W, Wx)


# ------------------------------------------------------------------ lists
def ref_state(D=6.0, H=8.0, seed=1):
    from oracle.oracle import OracleLayout

    c = cases.reference_uniaxial(D, H, seed=seed)
    return c, OracleLayout(c.R, c.K, c.grid())


def test_cell_distribution_is_stable_sort():
    c, o = ref_state()
    o.cell_distribution()
    assert o.last_status == 0
    g = c.grid()
    R = o.R.reshape(3, -1)
    ix = np.floor(R * np.float32(g.inv_a)).astype(np.int64)
    ci = ix[0] + ix[1] * g.Nx + ix[2] * g.Nx * g.Ny
    assert np.array_equal(o.CI, ci.astype(np.uint32))
    order = np.argsort(ci, kind="stable")
    assert np.array_equal(o.CIs[: o.N], ci[order].astype(np.uint32))
    assert np.array_equal(o.CIs[o.N:], order.astype(np.uint32))
    cnt = np.bincount(ci, minlength=g.C)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    assert np.array_equal(o.pnC[g.C:], cnt.astype(np.uint32))
    assert np.array_equal(o.pnC[: g.C], np.where(cnt > 0, start, 0).astype(np.uint32))


def brute_pairs(R, aacut):
    R = R.reshape(3, -1)
    N = R.shape[1]
    out = set()
    for i in range(N):
        d = (R[:, :] - R[:, i:i + 1]).astype(np.float32)
        rr = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
        for j in np.nonzero(rr < np.float32(aacut))[0]:
            if j != i:
                out.add((i, int(j)))
    return out


def test_il_construct_matches_brute_force():
    """h_CheckInteractions (md_check_test.cu:26-95): no pair within the cut is
    missing and no listed pair is beyond it."""
    c, o = ref_state()
    aacut = float(np.float32(c.mat.a_farcut) ** 2)
    o.cell_distribution()
    o.il_construct(aacut)
    assert o.last_status == 0
    got = set()
    IL = o.IL.reshape(o.N, o.K)
    for i in range(o.N):
        row = IL[i][IL[i] != 0xFFFFFFFF]
        assert len(set(row.tolist())) == len(row)
        got |= {(i, int(j)) for j in row}
    assert got == brute_pairs(o.R, aacut)


def test_il_row_order_own_cell_first():
    """Slot order: own cell first, then the +/-x neighbour cell, ids ascending."""
    from oracle.oracle import OracleLayout
    from rfd.reflayout import Grid

    a = np.float32(2.9)
    grid = Grid(float(a), float(np.float32(1) / a), 4, 4, 4)
    # particle 0 in cell (1,1,1) near its +x face; 1 in the +x cell; 2, 3 own cell
    P = np.array([[5.7, 4.35, 4.35], [6.0, 4.35, 4.35], [4.4, 4.35, 4.35], [5.0, 4.4, 4.3]],
                 np.float32)
    o = OracleLayout(np.ascontiguousarray(P.T).reshape(-1), 4, grid)
    o.cell_distribution()
    o.il_construct(float(np.float32(1.45) ** 2))
    assert o.IL[:4].tolist() == [2, 3, 1, 0xFFFFFFFF]


def test_reconstruct_first_free_slot_and_p0_quirk():
    from oracle.oracle import OracleLayout
    from rfd.reflayout import Grid

    a = np.float32(2.9)
    grid = Grid(float(a), float(np.float32(1) / a), 4, 4, 4)
    P = np.array([[4.35, 4.35, 4.35], [5.0, 4.35, 4.35], [4.35, 5.0, 4.35]], np.float32)
    for quirk, slot in ((True, 2), (False, 0)):
        o = OracleLayout(np.ascontiguousarray(P.T).reshape(-1), 4, grid)
        o.cell_distribution()
        o.IL[:4] = [0xFFFFFFFF, 1, 0xFFFFFFFF, 0xFFFFFFFF]
        o.ILtype[:4] = [9, 1, 9, 9]
        o.il_reconstruct_bpm(float(np.float32(1.45) ** 2), quirk_p0=quirk)
        row = o.IL[:4].tolist()
        assert row[1] == 1 and o.ILtype[1] == 1  # existing link kept, type untouched
        assert row[slot] == 2 and o.ILtype[slot] == 0


def test_bonds_construct_and_delete_far():
    c, o = ref_state()
    m = c.mat
    aacut = float(np.float32(m.a_farcut) ** 2)
    o.cell_distribution()
    o.il_construct(aacut)
    o.bonds_construct(m.b_dd)
    R = o.R.reshape(3, -1)
    IL = o.IL.reshape(o.N, o.K)
    ty = o.ILtype.reshape(o.N, o.K)
    for i in range(0, o.N, 7):
        for k in range(o.K):
            j = IL[i, k]
            if j == 0xFFFFFFFF:
                continue
            d = (R[:, j] - R[:, i]).astype(np.float32)
            rr = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
            assert ty[i, k] == (1 if rr < np.float32(m.b_dd) else 0)
            if ty[i, k]:
                assert o.iL[i * o.K + k] == pytest.approx(1 / math.sqrt(float(rr)), rel=1e-7)
    # pull one particle far away: all its links disappear on both sides
    o.R[0] += 10.0
    o.delete_far_links(m.aa_farcut)
    assert np.all(o.IL[: o.K] == 0xFFFFFFFF)
    assert not np.any(o.IL.reshape(o.N, o.K)[1:] == 0)


# ------------------------------------------------------- platens, loading
def test_platen_uniaxial_hand_sums():
    from oracle.oracle import OracleLayout
    from rfd.reflayout import Grid

    P = np.array([[0, 0, 9.9], [0.1, 0, 10.2], [0, 0, -0.1], [0, 0, 0.3], [9, 9, 9.9]],
                 np.float32)
    o = OracleLayout(np.ascontiguousarray(P.T).reshape(-1), 1, Grid(1, 1, 1, 1, 1))
    o.F[:] = np.array([0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 2.0, 3.0, -4.0, -1.5, 5.0], np.float32)
    cmp = Compression("uniaxial", (10, 10, 10), (0, 0, 5), 0.5, 0.01, -1e-3, 1.0, 0.1)
    s = o.platen_uniaxial(cmp, 100.0)
    # top band (z >= 9.5): particles 0 (F=2) and 1 (F=3); 1 is past Zt=10 -> clamped
    # bottom band (z <= 0.5): particles 2 (F=-4, past Zb=0 -> clamped) and 3
    # (F=-1.5, inside the band, not clamped); particle 4 is outside the radius
    assert s == (-5.5, 5.0)
    assert o.R[1 + 10] == 10.0 and o.F[1 + 10] == 0 and o.F[0 + 10] == 2.0
    assert o.R[2 + 10] == 0.0 and o.F[2 + 10] == 0
    assert o.R[3 + 10] == np.float32(0.3) and o.F[3 + 10] == np.float32(-1.5)


def test_fracture_detector_on_synthetic_trace():
    cmp = Compression("uniaxial", (10, 10, 10), (0, 0, 5), 0.5, 0.01, -1e-3, 1.0, 0.1)
    area = cmp.area
    trace = [0.5, 1.5, 3.0, 2.0, 0.9, 0.2]  # stress values
    for st in trace:
        cmp.sum_forces((-st * area, st * area))
        if cmp.fracture:
            break
    assert cmp.loading and cmp.fracture
    assert cmp.fracture_stress == pytest.approx(3.0)
    assert cmp.stress == pytest.approx(0.2)  # first value below 0.33*MinStress


def test_reference_fire_relaxes_quadratic_bowl():
    """Reference FIRE kernels + control (md_firerelaxation.cu:201-304) on
    F = -k (R - R0): converges to the analytic minimum R0 (its first
    iteration, at rest, takes the back-off branch: P = 0)."""
    from oracle.oracle import FireRefState, fire_relax_reference

    g = np.random.default_rng(7)
    N = 64
    R0 = g.uniform(0, 5, 3 * N).astype(np.float32)
    R = (R0 + g.normal(0, 0.1, 3 * N)).astype(np.float32)
    V = np.zeros_like(R)
    F = np.zeros_like(R)
    k = np.float32(2.0)

    def forces():
        F[:] = -k * (R - R0)

    f = FireRefState(0.05, 0.25)
    fire_relax_reference(F, V, R, N, 0.05, f, 1, forces)
    assert f.power == 0 and f.n_negative == 1 and f.dt == np.float32(0.05)
    fire_relax_reference(F, V, R, N, 0.05, f, 2999, forces)
    assert not f.stopped and f.steps == 3000
    assert f.dt == f.dtmax  # grown by 1.1 per positive step after the first 20
    assert np.abs(R - R0).max() < 1e-4


# ----------------------------------------------------------- native f64
def test_native_f64_matches_slot_layout_forces():
    """The FP64 edge-list oracle in reference mode agrees with the FP32 slot
    oracle on a random bonded pair sequence to FP32 accuracy; the pair moves
    by R += V dt between evaluations, the motion under which the edge-list
    form (previous direction at R - V dt, twist kept along n) restates the
    reference's stored rij and Mijn."""
    g = np.random.default_rng(11)
    m = mat()
    r = g.uniform(-1, 1, 3)
    r = r / np.linalg.norm(r) * 1.02
    p = Pair(tuple(r), L0=1.0)
    vs = g.standard_normal((4, 3)) * 1e-3
    p.set_v(0, vs[0]); p.set_v(1, vs[1]); p.set_w(0, vs[2]); p.set_w(1, vs[3])
    R = p.R.astype(np.float64); V = p.V.astype(np.float64); W = p.W.astype(np.float64)
    bi, bj = np.array([0], np.uint32), np.array([1], np.uint32)
    iL = np.array([1.0]); br = np.array([m.b_r]); sc = np.array([1e9]); tc = np.array([1e9])
    st = np.zeros(8); act = np.ones(1, np.uint8)
    F = np.zeros(6); M = np.zeros(6)
    for it in range(5):
        if it:
            p.R += p.V * np.float32(m.dt)
            R = p.R.astype(np.float64)
        p.bonds(m)
        lib().orc_native_bonds_f64(ptr(R), ptr(V), ptr(W), 2, ptr(bi), ptr(bj), ptr(iL),
                                   ptr(br), ptr(sc), ptr(tc), ptr(st), ptr(act), 1, m.dt,
                                   m.E, m.G, 1, ptr(F), ptr(M))
    np.testing.assert_allclose(F, p.F, rtol=2e-4, atol=1e-6 * np.abs(p.F).max())
    np.testing.assert_allclose(M, p.M, rtol=2e-4, atol=1e-6 * np.abs(p.M).max())


def test_native_f64_state_corotates_with_rigid_rotation():
    """Objectivity of the edge-list law: a bonded pair carrying shear, twist
    and bending state that turns rigidly (V = w x r, W = w, R += V dt) keeps
    its state fixed in the turning frame: after k evaluations O and Mt are the
    initial vectors turned by k atan(|w| dt) about w, mn is unchanged, and no
    new increment appears.  A previous direction taken at the wrong place (or
    not at all) leaves the state behind by the accumulated angle (0.02 rad)."""
    N = 2
    R = np.array([0.0, 1.0, 0.0, 0.0, 0.0, 0.0])  # SoA: x0 x1 y0 y1 z0 z1
    w = np.array([0.0, 0.0, 0.01])
    dt = 0.1
    th = np.arctan(np.linalg.norm(w) * dt)
    O0 = np.array([0.0, 2e-3, 1e-3])
    Mt0 = np.array([0.0, -1e-3, 3e-3])
    st = np.concatenate([O0, [5e-4], Mt0, [0.0]])
    act = np.ones(1, np.uint8)
    bi, bj = np.array([0], np.uint32), np.array([1], np.uint32)
    iL, br = np.array([1.0]), np.array([0.3])
    sc, tc = np.array([1e30]), np.array([1e30])
    F = np.zeros(6); M = np.zeros(6)
    W = np.tile(w[:, None], (1, 2)).reshape(-1)
    steps = 20
    for _ in range(steps):
        P = R.reshape(3, 2).T
        c = P.mean(0)
        V = np.cross(w, P - c).T.reshape(-1)
        lib().orc_native_bonds_f64(ptr(R), ptr(V), ptr(W), N, ptr(bi), ptr(bj), ptr(iL),
                                   ptr(br), ptr(sc), ptr(tc), ptr(st), ptr(act), 1, dt,
                                   30.0, 12.0, 0, ptr(F), ptr(M))
        R = R + V * dt
    a = steps * th
    Rz = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    np.testing.assert_allclose(st[0:3], Rz @ O0, rtol=0, atol=2e-5 * np.linalg.norm(O0))
    np.testing.assert_allclose(st[4:7], Rz @ Mt0, rtol=0, atol=2e-5 * np.linalg.norm(Mt0))
    assert st[3] == pytest.approx(5e-4, rel=1e-12)



# ------------------------------------------- stand-alone sum-up and break
def _bare_layout(N, K):
    from oracle.oracle import OracleLayout
    from rfd.reflayout import Grid

    return OracleLayout(np.zeros(3 * N, np.float32), K, Grid(1, 1, 1, 1, 1))


def test_sumup_known_answer():
    """md_sumupforcesdem.cu:19-60 on 2 particles x 2 slots with exactly
    representable loads: particle p sums its own row (slots p*K..p*K+K-1) of
    Fijn+Fijt into F and Mijn+Mijt+Mijadd into M, written SoA by component."""
    o = _bare_layout(2, 2)
    # slot s, component d: value 2^s * (d + 1) in Fijn, 3x that in Fijt, etc.
    for s in range(4):
        for d in range(3):
            o.slot["Fijn"][3 * s + d] = 2.0 ** s * (d + 1)
            o.slot["Fijt"][3 * s + d] = 3 * 2.0 ** s * (d + 1)
            o.slot["Mijn"][3 * s + d] = -(2.0 ** s)
            o.slot["Mijt"][3 * s + d] = 0.5 * d
            o.slot["Mijadd"][3 * s + d] = 100.0 * s
    o.sumup_forces()
    # particle 0: slots 0, 1 -> sum 2^s = 3; particle 1: slots 2, 3 -> 12
    for p, w, ss in ((0, 3.0, (0, 1)), (1, 12.0, (2, 3))):
        for d in range(3):
            assert o.F[p + 2 * d] == 4 * w * (d + 1)
            assert o.M[p + 2 * d] == -w + 2 * 0.5 * d + 100.0 * sum(ss)


def test_standalone_break_thresholds():
    """md_checkbreakconditiondem.cu:20-55: sigma = sign(Fn.n) |Fn| / A +
    |Mt| b / I, tau = |Ft| / A + |Mn| b / (2 I), I = pi b^4 / 4.  Loads just
    below and just above each threshold, and a compressive load whose stress
    counts negative."""
    b = 0.25
    A = float(PI_F) * b * b
    I = float(PI_F) * b ** 4 * 0.25
    sc, tc = 10.0, 20.0
    cases_ = [  # (Fn, Mt, Ft, Mn, broken)
        (0.999 * sc * A, 0, 0, 0, False),
        (1.001 * sc * A, 0, 0, 0, True),
        (0.5 * sc * A, 0.499 * sc * I / b, 0, 0, False),
        (0.5 * sc * A, 0.501 * sc * I / b, 0, 0, True),
        (-5 * sc * A, 5.9 * sc * I / b, 0, 0, False),   # compression offsets bending
        (0, 0, 0.999 * tc * A, 0, False),
        (0, 0, 1.001 * tc * A, 0, True),
        (0, 0, 0, 1.998 * tc * I / b, False),
        (0, 0, 0, 2.002 * tc * I / b, True),
    ]
    o = _bare_layout(1, len(cases_))
    n = np.array([0.6, 0.0, 0.8], np.float32)
    t = np.array([0.8, 0.0, -0.6], np.float32)
    for s, (fn, mt, ft, mn, _) in enumerate(cases_):
        o.ILtype[s] = 1
        o.slot["rij"][3 * s:3 * s + 3] = n
        o.slot["Fijn"][3 * s:3 * s + 3] = n * np.float32(fn)
        o.slot["Mijt"][3 * s:3 * s + 3] = t * np.float32(mt)
        o.slot["Fijt"][3 * s:3 * s + 3] = t * np.float32(ft)
        o.slot["Mijn"][3 * s:3 * s + 3] = n * np.float32(mn)
        o.slot["Oijt"][3 * s:3 * s + 3] = 1.0
    o.check_break(sc, tc, b)
    for s, c in enumerate(cases_):
        assert o.ILtype[s] == (0 if c[4] else 1), (s, c)
        assert np.all(o.slot["Oijt"][3 * s:3 * s + 3] == (0.0 if c[4] else 1.0))


def test_platen_brazil_hand_sums():
    """md_braziltest_simple2.cu:19-92: load axis y; inside the radius (x, z)
    and below Zcut, the top band sums F_y > 0 and brakes the in-plane force by
    0.45 F_y along the in-plane velocity; past Yt the particle is clamped."""
    P = np.array([[0, 9.9, 0], [0.1, 10.2, 0], [0, -0.1, 0], [0, 0.3, 0], [9, 9.9, 9]],
                 np.float32)
    N = len(P)
    R = np.ascontiguousarray(P.T).reshape(-1).copy()
    V = np.zeros(3 * N, np.float32)
    F = np.zeros(3 * N, np.float32)
    F[N + np.arange(N)] = [2.0, 3.0, -4.0, -1.5, 5.0]
    V[0], V[2 * N] = 3.0, 4.0          # particle 0 moves in-plane, |v| = 5
    F[0], F[2 * N] = 1.0, 1.0
    sums = (C.c_double * 2)()
    lib().orc_platen_brazil(ptr(R), ptr(V), ptr(F), N, 0.0, 0.0, 1.0, 10.0, 0.0, 9.5, 0.5,
                            -1e-3, 100.0, sums)
    assert (sums[0], sums[1]) == (-5.5, 5.0)
    # friction on particle 0: F -= 0.45 * 2 * v / |v|
    assert F[0] == pytest.approx(1.0 - 0.45 * 2 * 0.6, rel=1e-6)
    assert F[2 * N] == pytest.approx(1.0 - 0.45 * 2 * 0.8, rel=1e-6)
    assert R[N + 1] == 10.0 and V[N + 1] == np.float32(-1e-3) and F[N + 1] == 0
    assert R[N + 2] == 0.0 and F[N + 2] == 0
    assert R[N + 3] == np.float32(0.3) and F[N + 3] == np.float32(-1.5)
    assert F[N + 4] == 5.0  # outside the radius


def test_cut_restriction_moves_outside_particles():
    """md_parallelepipedrestriction.cu:73-158: a particle outside the box on
    any axis goes to the hidden point with V = F = 0; inside ones stay."""
    P = np.array([[0, 0, 0], [1.01, 0, 0], [0, -2.5, 0], [0, 0, 3.2], [0.99, -1.99, 2.99]],
                 np.float32)
    N = len(P)
    R = np.ascontiguousarray(P.T).reshape(-1).copy()
    V = np.ones(3 * N, np.float32)
    F = np.full(3 * N, 2.0, np.float32)
    c = np.zeros(3, np.float32); h = np.array([1, 2, 3], np.float32)
    hp = np.array([100, 200, 300], np.float32)
    lib().orc_cut_restriction(ptr(R), ptr(V), ptr(F), N, ptr(c), ptr(h), ptr(hp))
    out = [False, True, True, True, False]
    for i, o in enumerate(out):
        got = np.array([R[i], R[N + i], R[2 * N + i]])
        if o:
            assert np.all(got == hp) and V[i::N].max() == 0 and F[i::N].max() == 0
        else:
            assert np.all(got == P[i]) and V[i::N].min() == 1 and F[i::N].min() == 2


def test_platen_brazil_friction_opposes_sliding_on_both_platens():
    """md_braziltest_simple2.cu:19-92: a particle pressed on either platen
    (top: F_y > 0, bottom: F_y < 0) and sliding in the (x, z) plane gets a
    Coulomb friction force of 0.45 |F_y| against its sliding direction."""
    P = np.array([[0, 9.9, 0], [0, 0.1, 0]], np.float32)  # top band, bottom band
    N = len(P)
    R = np.ascontiguousarray(P.T).reshape(-1).copy()
    V = np.zeros(3 * N, np.float32)
    F = np.zeros(3 * N, np.float32)
    F[N + 0], F[N + 1] = 2.0, -4.0
    v = [(0.6, 0.8), (-0.6, 0.8)]  # unit in-plane velocities
    for i, (vx, vz) in enumerate(v):
        V[i], V[2 * N + i] = vx, vz
    sums = (C.c_double * 2)()
    lib().orc_platen_brazil(ptr(R), ptr(V), ptr(F), N, 0.0, 0.0, 1.0, 10.0, 0.0, 9.5, 0.5,
                            -1e-3, 100.0, sums)
    for i, (vx, vz) in enumerate(v):
        fy = 2.0 if i == 0 else -4.0
        fric = np.array([F[i], F[2 * N + i]])
        np.testing.assert_allclose(fric, -0.45 * abs(fy) * np.array([vx, vz]), rtol=1e-6)


def test_delete_far_links_threshold():
    """md_interaction_list_reconstructbpm.cu:141-186: a link whose squared
    length exceeds aa_farcut is removed from the row, one just inside stays."""
    aa = np.float32(2.25)
    for scale, gone in ((1.0 - 1e-3, False), (1.0 + 1e-3, True)):
        x = np.float32(math.sqrt(float(aa) * scale))
        R = np.array([0.0, x, 0, 0, 0, 0], np.float32)
        N, K = 2, 1
        IL = np.array([1, 0], np.uint32)
        ty = np.array([1, 1], np.uint8)
        st = [np.ones(6, np.float32) for _ in range(7)]
        lib().orc_delete_far_links(ptr(R), N, ptr(IL), ptr(ty), *[ptr(a) for a in st], K, aa)
        if gone:
            assert np.all(IL == 0xFFFFFFFF) and np.all(ty == 0)
            assert all(np.all(a == 0) for a in st)
        else:
            assert list(IL) == [1, 0] and np.all(ty == 1)
            assert all(np.all(a == 1) for a in st)


def test_fire_halfstep_moves_back_along_the_velocity():
    """md_firerelaxation.cu:58-70: the back-off after a negative power returns
    every particle half a step along its path, R <- R - V dt/2."""
    R = np.array([1.0, 2.0, 3.0, -1.0, 0.5, 4.0], np.float32)
    V = np.array([2.0, -4.0, 6.0, 8.0, 0.0, -2.0], np.float32)
    want = R - V * np.float32(0.25)
    lib().orc_fire_halfstep(ptr(V), ptr(R), 2, np.float32(0.25))
    np.testing.assert_array_equal(R, want)
