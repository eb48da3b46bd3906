"""Pins for the FP64 restatements of the native tier's explicit step
(oracle/native_step.py) and quasi-static load step (oracle/native_fire.py),
on the CPU.  The bond and contact laws they call are pinned in
test_oracle_pins.py and test_oracle_contacts.py; these tests hold what the
restatements add around them: symplectic Euler with per-particle mass and
solid-sphere inertia, the platen clamps and reaction sums, the loading control,
the cut box, the FIRE control of Bitzek et al. (2006) with the half-step
back-off, and a whole load step against the closed form of springs in series.
Each expected value is worked out from mechanics, not from the restatement."""
import math

import numpy as np
import pytest

from oracle.native_fire import FireState, OracleFireLoadStep
from oracle.native_step import OracleNativeStep
from rfd import native

E, NU, RHO, DT = 30.0, 0.25, 2.65, 1e-3
RAD, BR = 0.5, 0.4
MASS = RHO * 4.0 / 3.0 * math.pi * RAD ** 3


def pair(L, L0=1.0, **kw):
    R = np.array([[0.0, 0.0, 0.0], [L, 0.0, 0.0]])
    return OracleNativeStep(R, [RAD, RAD], [0], [1], [1.0 / L0], [BR], [1e9], [1e9], E=E, nu=NU,
                            rho=RHO, mu=0.45, mur=0.05, e=0.5, dt=DT, **kw)


def loose(R, **kw):
    """Unbonded particles at R."""
    return OracleNativeStep(np.asarray(R, np.float64), [RAD] * len(R), [], [], [], [], [], [],
                            E=E, nu=NU, rho=RHO, mu=0.45, mur=0.05, e=0.5, dt=DT, **kw)


def loading(**kw):
    L = {name: 0 for name, _ in native.Loading._fields_}
    for k in ("box_c", "box_h", "hidden"):
        L[k] = [0.0, 0.0, 0.0]
    L.update(kw)
    return L


def test_explicit_step_of_a_stretched_pair():
    """One step from rest of a bond stretched by eps: each end is pulled by
    E A eps (A = pi b^2), then V = F dt / m and R += V dt (symplectic Euler,
    md_calculateincrementsdem.cu:19-49) with m = rho 4/3 pi r^3."""
    eps = 1e-3
    o = pair(1.0 + eps)
    o.step()
    f = E * math.pi * BR * BR * eps
    np.testing.assert_allclose(o.F[0], [f, 0, 0], rtol=1e-12, atol=0)
    np.testing.assert_allclose(o.F[1], [-f, 0, 0], rtol=1e-12, atol=0)
    v = f * DT / MASS
    np.testing.assert_allclose(o.V[0], [v, 0, 0], rtol=1e-12, atol=0)
    np.testing.assert_allclose(o.R[0], [v * DT, 0, 0], rtol=1e-12, atol=0)
    np.testing.assert_allclose(o.R[1], [1.0 + eps - v * DT, 0, 0], rtol=1e-15, atol=0)
    assert np.all(o.W == 0) and np.all(o.M == 0)


def test_spin_increment_uses_solid_sphere_inertia():
    """j spins about the bond axis at w: after one step the twisting moment on
    i is G (pi b^4 / 2) (w dt) / L0, and W_i = M dt / I with I = 2/5 m r^2."""
    w = 2e-3
    o = pair(1.0, W=[[0.0, 0.0, 0.0], [w, 0.0, 0.0]])
    o.step()
    G = E / (2.0 * (1.0 + NU))
    Mx = G * (math.pi * BR ** 4 / 2.0) * w * DT / 1.0
    np.testing.assert_allclose(o.M[0], [Mx, 0, 0], rtol=1e-12, atol=1e-30)
    inertia = 0.4 * MASS * RAD * RAD
    np.testing.assert_allclose(o.W[0], [Mx * DT / inertia, 0, 0], rtol=1e-12, atol=1e-30)
    np.testing.assert_allclose(o.W[1], [w - Mx * DT / inertia, 0, 0], rtol=1e-12)


def test_uniaxial_platens_clamp_and_sum():
    """md_uniaxialcompression3_simple.cu:19-76 on FP64 positions: inside the
    radius, the top band sums F_z > 0 and the bottom band F_z < 0; a particle
    past a platen is put on it, takes the platen's velocity and loses F_z."""
    P = np.array([[0, 0, 9.9], [0, 0, 10.2], [0, 0, -0.1], [0, 0, 0.3], [20, 0, 9.9],
                  [0, 0, 9.8]], np.float64)
    F = np.zeros_like(P)
    F[:, 2] = [2.0, 3.0, -4.0, -1.5, 5.0, -2.0]
    V = np.ones_like(P)
    o = loose(P)
    L = loading(mode=1, cx=0.0, cy=0.0, RR=100.0, zcut=1e30, top=10.0, bottom=0.0,
                top_r=9.5, bottom_r=0.5, v=-0.01)
    R = P.copy()
    sb, st = o._platen(R, V, F, L=L)
    assert (sb, st) == (-5.5, 5.0)
    assert R[1, 2] == 10.0 and V[1, 2] == -0.01 and F[1, 2] == 0.0
    assert R[2, 2] == 0.0 and V[2, 2] == 0.0 and F[2, 2] == 0.0
    assert R[0, 2] == 9.9 and F[0, 2] == 2.0 and V[0, 2] == 1.0
    assert F[4, 2] == 5.0 and F[5, 2] == -2.0  # outside the radius; pulled, not pressed
    assert np.all(V[:, :2] == 1.0) and np.all(F[:, :2] == 0.0)


def test_brazil_platens_friction_opposes_sliding():
    """md_braziltest_simple2.cu:19-92: load axis y; a particle pressed on
    either platen and sliding in the (x, z) plane is braked by 0.45 |F_y|
    (the reference's 0.45f) against its sliding direction."""
    P = np.array([[0, 9.9, 0], [0, 0.1, 0]], np.float64)
    V = np.array([[0.6, 0.0, 0.8], [-0.6, 0.0, 0.8]])
    F = np.array([[0.0, 2.0, 0.0], [0.0, -4.0, 0.0]])
    o = loose(P)
    L = loading(mode=2, cx=0.0, cz=0.0, RR=1.0, zcut=100.0, top=10.0, bottom=0.0,
                top_r=9.5, bottom_r=0.5)
    sb, st = o._platen(P.copy(), V.copy(), F, L=L)
    assert (sb, st) == (-4.0, 2.0)
    mu = float(np.float32(0.45))
    np.testing.assert_allclose(F[0, [0, 2]], -mu * 2.0 * V[0, [0, 2]], rtol=1e-15)
    np.testing.assert_allclose(F[1, [0, 2]], -mu * 4.0 * V[1, [0, 2]], rtol=1e-15)


def test_loading_control_trace():
    """SumForcesULoading (md_results.cu:34-74) with CompressU
    (md_data_types.h:223-244) on a hand-made trace of platen sums: stress is
    the mean of |bottom| and |top| over the area; the platen runs at 10 v_load
    until the first contact, then at v_load; loading starts above min_stress;
    the peak and its strain are kept; fracture once loading and the stress
    falls below zero_stress."""
    o = pair(1.0)
    o.L = loading(mode=1, size0=10.0, size=10.0, top=10.0, band=0.5, v_load=-0.01,
                  inv_area=0.5, min_stress=0.1, zero_stress=0.05)
    trace = [(0.0, 0.0), (-0.1, 0.1), (-0.4, 0.4), (-0.6, 0.6), (-0.3, 0.2), (-0.08, 0.08)]
    want_stress = [0.0, 0.05, 0.2, 0.3, 0.125, 0.04]
    size = 10.0
    peak, peak_strain = 0.0, 0.0
    for k, ((s0, s1), sw) in enumerate(zip(trace, want_stress)):
        o.control(s0, s1)
        L = o.L
        assert L["stress"] == pytest.approx(sw, rel=1e-15)
        assert L["loading"] == (1 if k >= 2 else 0)
        if sw > peak:
            peak, peak_strain = sw, (10.0 - size) / 10.0
        assert L["fracture_stress"] == pytest.approx(peak, rel=1e-15)
        assert L["fracture_strain"] == pytest.approx(peak_strain, rel=1e-12, abs=1e-15)
        v = -0.1 if sw < 1e-9 else -0.01
        assert L["v"] == v
        size += v * DT
        assert L["size"] == pytest.approx(size, rel=1e-15)
        assert L["top"] == pytest.approx(size, rel=1e-15)
        assert L["top_r"] == pytest.approx(size - 0.5, rel=1e-15)
        assert L["fracture"] == (1 if k == 5 else 0)
    assert L["steps"] == len(trace)


def test_cut_box_hides_escaped_particles():
    """md_parallelepipedrestriction.cu:73-158 after the update: a particle
    outside the box on any axis goes to the hidden point with V = F = 0."""
    R = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 2.5, 0.0]])
    o = loose(R, V=np.ones((3, 3)),
              loading=loading(mode=0, cut=1, box_c=[0.0, 0.0, 0.0], box_h=[1.5, 2.0, 1.5],
                              hidden=[100.0, 200.0, 300.0]))
    o.F = np.full((3, 3), 2.0)
    o.M = np.zeros((3, 3))
    o.particle_pass()
    assert np.all(o.R[2] == [100.0, 200.0, 300.0]) and np.all(o.V[2] == 0) and np.all(o.F[2] == 0)
    for i in (0, 1):
        v = 1.0 + 2.0 * DT / MASS
        np.testing.assert_allclose(o.V[i], [v, v, v], rtol=1e-15)
        np.testing.assert_allclose(o.R[i], R[i] + v * DT, rtol=1e-15)


def _fire():
    fp = native.fire_params(E, RAD, DT, rel_tol=1e-8)
    o = pair(1.0)
    return OracleFireLoadStep(o, fp), fp


def test_fire_control_rules():
    """FIRE (Bitzek et al. 2006): while the power P = F.V + M.W is positive,
    mixing uses cV = alpha |V| / |F| (cW = alpha |W| / |M|); after more than
    n_delay positive steps dt grows by f_inc (capped at dt_max) and alpha
    shrinks by f_alpha; on P <= 0 the step backs off, dt shrinks by f_dec
    (floored at dt0 / 1000) and alpha returns to alpha0.  Converged once
    max |F_i| and max |M_i| / r_i are below ftol."""
    s, fp = _fire()
    f = FireState(fp)
    big = 100.0 * fp.ftol
    up = [1.0, 4.0, 9.0, 16.0, 25.0, big * big, big * big, 0.0, 0.0]
    s._control(f, up)
    assert not f.done and not f.backoff and f.n_pos == 1 and f.dt == fp.dt0
    assert f.cV == pytest.approx(fp.alpha0 * 3.0 / 2.0, rel=1e-15)
    assert f.cW == pytest.approx(fp.alpha0 * 5.0 / 4.0, rel=1e-15)
    for _ in range(fp.n_delay):
        s._control(f, up)
    assert f.n_pos == fp.n_delay + 1
    assert f.dt == pytest.approx(fp.dt0 * fp.f_inc, rel=1e-15)
    assert f.alpha == pytest.approx(fp.alpha0 * fp.f_alpha, rel=1e-15)
    for _ in range(200):
        s._control(f, up)
    assert f.dt == fp.dt_max
    down = [-1.0] + up[1:]
    s._control(f, down)
    assert f.backoff and f.n_pos == 0 and f.alpha == fp.alpha0
    assert f.dt == pytest.approx(fp.dt_max * fp.f_dec, rel=1e-15)
    assert f.dt_last == fp.dt_max  # the back-off undoes half of the last move
    for _ in range(100):
        s._control(f, down)
    assert f.dt == pytest.approx(1e-3 * fp.dt0, rel=1e-15)
    small = 0.5 * fp.ftol
    s._control(f, [1.0, 1.0, 1.0, 1.0, 1.0, small * small, small * small, -3.0, 3.0])
    assert f.done and f.converged
    assert (f.sum_bottom, f.sum_top) == (-3.0, 3.0)


def test_fire_update_backoff_and_kick():
    """The FIRE move: after a back-off the particles go back half of the last
    step along their FIRE velocity, which is then zeroed; otherwise it is mixed
    toward F, V <- (1 - alpha) V + cV F.  Then the kick V += F dt / m and the
    move R += V dt (the spins likewise with I = 2/5 m r^2)."""
    s, fp = _fire()
    o = s.o
    R0 = o.R.copy()
    o.F = np.array([[1.0, 0.0, 0.0], [0.0, 2.0, 0.0]])
    o.M = np.array([[0.0, 0.0, 3.0], [0.0, 0.0, 0.0]])
    s.Vf = np.array([[0.5, 0.0, 0.0], [0.0, -1.0, 0.0]])
    o.W = np.zeros((2, 3))
    f = FireState(fp)
    f.backoff, f.dt_last, f.dt = True, 4e-3, 2e-3
    s._update(f)
    inertia = 0.4 * MASS * RAD * RAD
    Vk = o.F * (2e-3 / MASS)
    np.testing.assert_allclose(s.Vf, Vk, rtol=1e-15)
    want = R0 - np.array([[0.5, 0, 0], [0, -1.0, 0]]) * 2e-3 + Vk * 2e-3
    np.testing.assert_allclose(o.R, want, rtol=1e-15, atol=1e-18)
    np.testing.assert_allclose(o.W, o.M * (2e-3 / inertia), rtol=1e-15)
    # mixing branch
    s.Vf = np.array([[0.5, 0.0, 0.0], [0.0, -1.0, 0.0]])
    o.W = np.zeros((2, 3))
    f.backoff, f.alpha, f.cV, f.cW = False, 0.1, 0.25, 0.0
    s._update(f)
    mixed = np.array([[0.5, 0, 0], [0, -1.0, 0]]) * 0.9 + o.F * 0.25
    np.testing.assert_allclose(s.Vf, mixed + Vk, rtol=1e-15)


def test_load_step_of_a_bonded_chain_reaches_series_springs():
    """n spheres on the z axis at spacing 1, bonded to their neighbours,
    between the uniaxial platens: one load step of strain eps relaxes to equal
    spacing 1 + eps, and both platens carry the force of every bond in the
    series, E pi b^2 |eps|."""
    n, eps = 8, -2e-4
    R = np.zeros((n, 3))
    R[:, 0] = R[:, 1] = 5.0
    R[:, 2] = 5.0 + np.arange(n, dtype=np.float64)
    bi = np.arange(n - 1)
    L = loading(mode=1, cx=5.0, cy=5.0, RR=4.0, zcut=1e30, bottom=5.0, top=5.0 + (n - 1),
                band=0.5, top_r=5.0 + (n - 1) - 0.5, bottom_r=5.25, size=n - 1.0,
                size0=n - 1.0, inv_area=1.0)
    o = OracleNativeStep(R, [RAD] * n, bi, bi + 1, np.ones(n - 1), [BR] * (n - 1),
                         [1e9] * (n - 1), [1e9] * (n - 1), E=E, nu=NU, rho=RHO, mu=0.45,
                         mur=0.05, e=0.5, dt=DT, loading=L)
    fp = native.fire_params(E, RAD, DT, rel_tol=1e-10, max_iter=200000)
    r = OracleFireLoadStep(o, fp).load_step(eps)
    assert r["converged"] and r["breaks"] == 0
    Fbond = E * math.pi * BR * BR * abs(eps)
    assert r["sum_top"] == pytest.approx(Fbond, rel=1e-4)
    assert r["sum_bottom"] == pytest.approx(-Fbond, rel=1e-4)
    np.testing.assert_allclose(o.R[:, 2], 5.0 + np.arange(n) * (1.0 + eps), rtol=0, atol=1e-8)
    np.testing.assert_allclose(o.R[:, :2], 5.0, rtol=0, atol=1e-12)


def test_fire_measure_power_and_norms_include_rotations():
    """The FIRE measure of the standard FIRE with rotations: P = F.V + M.W,
    |F|^2 + ..., |V|^2, |M|^2, |W|^2, max |F_i|^2, max (|M_i| / r_i)^2, each
    from the FIRE velocity and the spins."""
    s, fp = _fire()
    o = s.o
    o.L = loading(mode=1, RR=0.0, zcut=1e30)  # no particle inside the platens' radius
    o.F = np.array([[1.0, 2.0, 0.0], [0.0, -1.0, 3.0]])
    o.M = np.array([[0.0, 0.5, 0.0], [0.25, 0.0, 0.0]])
    s.Vf = np.array([[2.0, 0.0, 1.0], [1.0, 1.0, 1.0]])
    o.W = np.array([[0.0, 4.0, 0.0], [-2.0, 0.0, 0.0]])
    v = s._measure(FireState(fp))
    assert v[0] == pytest.approx((2.0 + 2.0) + (2.0 - 0.5), rel=1e-15)
    assert v[1:5] == pytest.approx([5.0 + 10.0, 5.0 + 3.0, 0.25 + 0.0625, 16.0 + 4.0])
    assert v[5] == 10.0
    assert v[6] == pytest.approx(0.25 / (RAD * RAD), rel=1e-15)
    assert v[7:] == [0.0, 0.0]


def test_fire_converges_only_when_moments_are_small_too():
    """Convergence needs max |F_i| < ftol and max |M_i| / r_i < ftol."""
    s, fp = _fire()
    f = FireState(fp)
    small, big = 0.5 * fp.ftol, 2.0 * fp.ftol
    s._control(f, [1.0, 1.0, 1.0, 1.0, 1.0, small * small, big * big, 0.0, 0.0])
    assert not f.done and not f.converged
    s._control(f, [1.0, 1.0, 1.0, 1.0, 1.0, small * small, small * small, 0.0, 0.0])
    assert f.done and f.converged


@pytest.mark.parametrize("refmode", [0, 1])
def test_broken_bond_contact_reach(refmode):
    """A bond made at L0 = 0.97 between spheres of radius 0.5 (overlapping by
    0.03) and then broken: in the physical mode its contact reaches min(r_i +
    r_j, L0), so at L = 0.98 it carries no load and at 0.95 the Hertz force of
    an overlap 0.02; in the reference mode the contact reaches r_i + r_j =
    1.0, as md_calculateforcesdem3.cu:338-422 does for every unbonded slot."""
    Estar = 0.5 * E / (1.0 - NU * NU)
    Rs = RAD / 2.0
    for L, d_phys, d_ref in ((0.98, 0.0, 0.02), (0.95, 0.02, 0.05)):
        o = pair(L, L0=0.97, refmode=refmode)
        o.act[:] = 0
        o.forces(quasi=True)
        d = d_ref if refmode else d_phys
        want = (4.0 / 3.0) * Estar * math.sqrt(Rs) * d ** 1.5
        assert o.F[0, 0] == pytest.approx(-want, rel=1e-12, abs=1e-300)
        assert o.F[1, 0] == pytest.approx(want, rel=1e-12, abs=1e-300)


def test_break_rounds_restart_fire_from_rest(monkeypatch):
    """After bonds break at equilibrium the step relaxes again from rest: each
    round starts with no velocity, the initial step dt0, mixing alpha0 and a
    zero iteration count (so the step's iterations are the rounds' sum)."""
    n, eps = 6, -2e-4
    R = np.zeros((n, 3))
    R[:, 0] = R[:, 1] = 5.0
    R[:, 2] = 5.0 + np.arange(n, dtype=np.float64)
    bi = np.arange(n - 1)
    # the middle bond's tensile strength is below the compressive E eps of the
    # step (negative), so it breaks at the first judgement, then meets the
    # contact law
    sc = np.full(n - 1, 1e9)
    sc[2] = -2.0 * E * abs(eps)
    L = loading(mode=1, cx=5.0, cy=5.0, RR=4.0, zcut=1e30, bottom=5.0, top=5.0 + (n - 1),
                band=0.5, top_r=5.0 + (n - 1) - 0.5, bottom_r=5.25, size=n - 1.0,
                size0=n - 1.0, inv_area=1.0)
    o = OracleNativeStep(R, [RAD] * n, bi, bi + 1, np.ones(n - 1), [BR] * (n - 1), sc,
                         [1e9] * (n - 1), E=E, nu=NU, rho=RHO, mu=0.45, mur=0.05, e=0.5,
                         dt=DT, loading=L)
    fp = native.fire_params(E, RAD, DT, rel_tol=1e-8, max_iter=200000)
    s = OracleFireLoadStep(o, fp)
    seen, counts = [], []
    relax = s._relax

    def spy(f, forces_current):
        seen.append((f.iter, f.dt, f.alpha, float(np.abs(s.Vf).max()),
                     float(np.abs(o.V).max()), float(np.abs(o.W).max())))
        out = relax(f, forces_current)
        counts.append(out.iter)
        return out

    monkeypatch.setattr(s, "_relax", spy)
    r = s.load_step(eps)
    assert r["converged"] and r["breaks"] == 1 and r["break_rounds"] == 1
    assert len(seen) == 2
    for it, dt, alpha, vf, v, w in seen:
        assert (it, dt, alpha) == (0, fp.dt0, fp.alpha0)
        assert vf == 0.0 and w == 0.0
    assert seen[1][4] == 0.0  # the first round starts from the platen's move
    assert r["iterations"] == sum(counts)
