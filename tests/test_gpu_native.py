"""GPU parity of the native tier (edge-list bonds, FP64 positions).

* reference mode vs the FP32 slot-layout oracle on the reference's own
  monodisperse material: one force evaluation from the same state;
* physical mode vs the FP64 edge-list oracle on a polydisperse specimen:
  bond forces and moments over several evaluations;
* invariants: Newton's third law over the whole specimen, break bookkeeping;
* a loading run against the reference-layout tier.
Tolerances are stated per test (relative to the largest component).
"""
import numpy as np
import pytest

from oracle.oracle import OracleLayout, lib as olib, native_contacts, ptr
from rfd import cases, native, native_cases
from rfd.reflayout import UniaxialDriver

pytestmark = pytest.mark.gpu


def _ref_setup(case):
    o = OracleLayout(case.R, case.K, case.grid())
    m = case.mat
    aacut = float(np.float32(m.a_farcut) ** 2)
    o.cell_distribution()
    o.il_construct(aacut)
    o.bonds_construct(m.b_dd)
    o.delete_far_links(m.aa_farcut)
    return o


def _soa3(a, N):
    return a.reshape(3, N).T


def test_refmode_forces_match_slot_oracle(ctx):
    """FP32 slot oracle vs native (refmode=1) from the same state; the native
    stretch is formed in FP64, so the tolerance is 2e-4 of max|F| (the FP32
    (L - L0)/L0 of the reference loses ~6e-8/strain relative)."""
    case = cases.reference_uniaxial(8.0, 12.0, seed=13)
    o = _ref_setup(case)
    N, m = case.N, case.mat
    g = np.random.default_rng(5)
    dR = (g.standard_normal(3 * N) * 2e-3).astype(np.float32)
    o.R += dR
    o.V[:] = (g.standard_normal(3 * N) * 1e-3).astype(np.float32)
    o.W[:] = (g.standard_normal(3 * N) * 2e-3).astype(np.float32)
    h = o.host()
    b, st, fl = native.bonds_from_slots(h, N, case.K, m)
    R64 = _soa3(o.R.astype(np.float64), N)
    p = native.params_from_material(m, refmode=1)
    s = native.NativeSystem(ctx, p, R64, np.full(N, m.p_r, np.float32))
    s.set_bonds(b)
    s.set_bond_state(st, fl)
    s.set_velocities(o.V, o.W)
    s.compute_forces()
    out = s.state()
    o.zero_fm()
    o.forces_bpm(m)
    o.forces_contact(m)
    for k in ("F", "M"):
        scale = np.abs(getattr(o, k)).max()
        assert scale > 0
        np.testing.assert_allclose(out[k], getattr(o, k), rtol=0, atol=2e-4 * scale, err_msg=k)
    # the same bonds break, and the surviving bonds' new state matches the
    # owner slot's state (pairs mapped by (i, j))
    st2, fl2 = s.bond_state()
    b2, st_ref, _ = native.bonds_from_slots(o.host(), N, case.K, m)
    key = lambda bi, bj: bi.astype(np.int64) * N + bj
    alive = np.isin(key(b.bi, b.bj), key(b2.bi, b2.bj))
    assert np.array_equal(alive, fl2 == 1)
    assert 0 < alive.sum() < len(alive)  # some bonds broke in this evaluation
    np.testing.assert_allclose(st2[alive][:, 3:], st_ref[:, 3:], rtol=0,
                               atol=2e-4 * np.abs(st_ref[:, 3:]).max())
    s.close()


def _poly_system(ctx, refmode=0, D=6.0, H=8.0, seed=3, strong=True):
    strength = (1e12, 1e12) if strong else (20e6, 40e6)
    nc = native_cases.uniaxial_cylinder(ctx, D, H, seed=seed, refmode=refmode, strength=strength)
    return nc


def test_physical_mode_matches_f64_oracle(ctx):
    """Polydisperse bonds, per-bond radius and the BASELINE strengths
    N(20,4) / N(40,8) MPa: native FP32 bond law (FP64 stretch, FP32 criterion
    rechecked in FP64 near the threshold) vs the FP64 oracle over 6
    evaluations at fixed positions, the state carried from one to the next;
    the intact/broken flag of every bond equal after each evaluation (bonds
    break from the second on), forces and moments within 1e-4 of the largest
    (bonds breaking in an evaluation meet the contact law in it, as in the
    oracle's restatement of the whole evaluation).  Velocities are large
    enough (|dV| dt ~ 2e-4 L) that the previous bond direction taken at
    R - V dt turns the carried state by a visible angle, and the shear and
    bending they drive break bonds evaluation by evaluation."""
    from test_gpu_native_oracle import oracle_of

    nc = native_cases.uniaxial_cylinder(ctx, 6.0, 8.0, seed=3, refmode=0,
                                        strength=(20e6, 40e6))
    s = nc.sys
    N = s.N
    g = np.random.default_rng(9)
    V = (g.standard_normal(3 * N) * 5e-3).astype(np.float32)
    W = (g.standard_normal(3 * N) * 1e-2).astype(np.float32)
    s.set_velocities(V, W)
    o = oracle_of(s)
    broken = []
    for it in range(6):
        s.compute_forces()
        o.forces()
        fl = s.bond_state()[1]
        assert np.array_equal(fl == 1, o.act == 1), f"bond flags differ at evaluation {it}"
        broken.append(int(np.count_nonzero(o.act == 0)))
        out = s.state64()
        for k, ref in (("F", o.F), ("M", o.M)):
            scale = np.abs(ref).max()
            np.testing.assert_allclose(out[k], ref, rtol=0, atol=1e-4 * scale,
                                       err_msg=f"{k} at evaluation {it}")
    print(f"broken bonds after each evaluation: {broken} of {s.bonds.n}", flush=True)
    assert 0 < broken[0] < broken[-1] < s.bonds.n


@pytest.mark.parametrize("refmode", [0, 1])
def test_contacts_of_broken_bonds_and_new_pairs_match_oracle(ctx, refmode):
    """Every bond broken, the specimen squeezed along z and jittered so that
    bonded pairs come closer than their rest length and never-bonded pairs
    come into contact, random V and W: one force
    evaluation against the FP64 native contact law of the oracle, per
    directed pair (broken bonds and never-bonded pairs within ri + rj).  A
    broken bond overlaps from min(ri + rj, L0) in the physical mode and from
    ri + rj in the reference mode.  Tolerance 1e-4 of max |F| and max |M|."""
    from scipy.spatial import cKDTree

    nc = native_cases.uniaxial_cylinder(ctx, 6.0, 8.0, seed=21, refmode=refmode)
    s = nc.sys
    N = s.N
    R0 = _soa3(s.state()["R"], N)
    b = s.bonds
    L0 = np.linalg.norm(R0[b.bj] - R0[b.bi], axis=1)
    iL = (1.0 / L0).astype(np.float32).astype(np.float64)
    g = np.random.default_rng(23)
    # squeeze along z: second lattice neighbours along z (sqrt(2) apart) come
    # within ri + rj, bonded pairs along z come closer than L0
    c = R0.mean(0)
    R = c + (R0 - c) * np.array([1.0, 1.0, 0.82]) + g.standard_normal((N, 3)) * 0.02
    s.set_positions(R)
    s.set_bond_state(np.zeros((b.n, 12), np.float32), np.zeros(b.n, np.uint8))
    V = (g.standard_normal((N, 3)) * 2e-2).astype(np.float32)
    W = (g.standard_normal((N, 3)) * 4e-2).astype(np.float32)
    s.set_velocities(V.T.reshape(-1), W.T.reshape(-1))
    s.compute_forces()
    out = s.state()
    rad = s.radius.astype(np.float64)
    sumr = rad[b.bi] + rad[b.bj]
    reach_b = sumr if refmode else np.minimum(sumr, 1.0 / iL)
    bonded = set(zip(b.bi.tolist(), b.bj.tolist()))
    cand = cKDTree(R).query_pairs(1.3, output_type="ndarray")
    d = np.linalg.norm(R[cand[:, 1]] - R[cand[:, 0]], axis=1)
    keep = np.array([(int(i), int(j)) not in bonded for i, j in cand]) & \
        (d < rad[cand[:, 0]] + rad[cand[:, 1]])
    newp = cand[keep]
    pairs = np.concatenate([np.stack([b.bi, b.bj], 1), np.stack([b.bj, b.bi], 1),
                            newp, newp[:, ::-1]]).astype(np.uint32)
    reach = np.concatenate([reach_b, reach_b, rad[newp[:, 0]] + rad[newp[:, 1]],
                            rad[newp[:, 0]] + rad[newp[:, 1]]])
    m = nc.mat
    O = np.zeros((len(pairs), 3))
    Fo, Mo = native_contacts(R, V, W, rad, pairs, reach, O, m.dt, m.hm_E, m.hm_G, m.p_A,
                             m.p_mu, m.p_mur, m.rho * 4.0 / 3.0 * np.pi)
    n_touch = int(np.count_nonzero(np.linalg.norm(R[pairs[:, 1]] - R[pairs[:, 0]], axis=1)
                                   < reach))
    assert n_touch > 50 and len(newp) > 0
    for k, ref in (("F", Fo), ("M", Mo)):
        got = _soa3(out[k], N)
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-4 * np.abs(ref).max(),
                                   err_msg=k)


def test_newton3_whole_specimen(ctx):
    nc = _poly_system(ctx, refmode=0, seed=4)
    s = nc.sys
    N = s.N
    g = np.random.default_rng(1)
    s.set_velocities((g.standard_normal(3 * N) * 1e-3).astype(np.float32),
                     (g.standard_normal(3 * N) * 1e-3).astype(np.float32))
    for _ in range(3):
        s.compute_forces()
    F = _soa3(s.state()["F"].astype(np.float64), N)
    total = F.sum(0)
    assert np.abs(total).max() <= 1e-5 * np.abs(F).sum(0).max()


def test_weak_bonds_break_consistently(ctx):
    nc = native_cases.uniaxial_cylinder(ctx, 6.0, 8.0, seed=6, strength=(2e6, 4e6))
    s = nc.sys
    N = s.N
    g = np.random.default_rng(2)
    s.set_velocities((g.standard_normal(3 * N) * 5e-2).astype(np.float32), None)
    c0 = s.bond_counts()
    s.step(30)
    c1 = s.bond_counts()
    assert c1[1] > 0 and c1[0] < c0[0]
    assert sum(c1) == sum(c0)
    out = s.state()
    assert np.isfinite(out["R"]).all() and np.isfinite(out["F"]).all()
    assert ctx.flags() == 0


def test_native_loading_tracks_reference_tier(ctx):
    """Same monodisperse specimen, material and rest lengths, reference-layout
    tier vs native tier in reference mode, 450 explicit steps of fast uniaxial
    loading that engages the platens: once the reaction exceeds 5 % of the
    peak, the platen stress histories agree to 3 % of the peak, and the peaks
    agree to 3 % (FP64 vs FP32 positions and a different summation order make
    the trajectories drift slowly)."""
    cr = cases.reference_uniaxial(8.0, 14.0, seed=21)
    cn = cases.reference_uniaxial(8.0, 14.0, seed=21)
    for c in (cr, cn):
        c.cmp.v_load *= 150
    dr = UniaxialDriver(ctx, cr.R, cr.mat, cr.box, cr.cmp, cr.K)
    dr.setup()
    h = dr.s.host()
    N = cr.N
    b, st, fl = native.bonds_from_slots(h, N, cr.K, cr.mat)
    p = native.params_from_material(cn.mat, refmode=1, rebuild_every=1)
    s = native.NativeSystem(ctx, p, _soa3(h["R"].astype(np.float64), N),
                            np.full(N, cn.mat.p_r, np.float32))
    s.set_bonds(b)
    s.set_bond_state(st, fl)
    zcut = float(np.float32(cn.box.center[2]) + np.float32(cn.box.spacesized2[2]))
    s.set_loading(native.loading_from_compression(cn.cmp, zcut, cn.box))
    steps = 450
    for _ in range(steps):
        dr.step()
    s.step(steps)
    hist_n = s.history()
    hist_r = np.array(cr.cmp.history)
    assert hist_n.shape == hist_r.shape
    np.testing.assert_allclose(hist_n[:, 0], hist_r[:, 0], rtol=0, atol=1e-9)
    mag_r = np.abs(hist_r[:, 1:]).max(1)
    peak = mag_r.max()
    assert peak > 1e-4  # the platens carry load (> 100 kPa)
    # the rising branch up to the reference's peak (fracture afterwards is
    # chaotic: bond-break order amplifies rounding differences)
    on = (mag_r > 0.05 * peak) & (np.arange(len(mag_r)) <= int(np.argmax(mag_r)))
    assert on.sum() > 30
    np.testing.assert_allclose(hist_n[on, 1:], hist_r[on, 1:], rtol=0, atol=0.03 * peak)
    assert np.abs(hist_n[:, 1:]).max() == pytest.approx(peak, rel=0.03)
