"""Whole explicit steps of the native tier against the CPU FP64 restatement
(oracle/native_step.py: the oracle's FP64 bond and contact laws, platens,
integration, cut box and loading control), and the break decision at the
threshold against the FP64 bond law.

* the 2k-sphere Brazilian disc (random packing at solid fraction 0.6, bonds at
  gap < 0.05 r_mean, strengths N(20,4) / N(40,8) MPa) on the FP64 path,
  step by step: positions to 1e-9 mm, forces and moments to 1e-6 of the
  largest (the API hands them out as FP32 copies), the intact/broken flag of
  every bond exactly, the platen reactions to 1e-9 relative.  The platens
  close at 20 m/s, so that bonds break at the platens within the run and the
  broken bonds and new pairs meet the contact law;
* the same on a lattice cylinder under uniaxial platens with the cut box;
* the disc and a batch member on the FP32 path against the same FP64
  restatement, at tolerances stated per test;
* bonds whose strength lies within FP32 rounding of their stress.
Every run is seeded, so each test sees the same trajectory on every run.
"""
import numpy as np
import pytest

from oracle.native_step import OracleNativeStep
from oracle.oracle import lib as olib, ptr
from rfd import native_cases
from rfd.params import VELOCITY

pytestmark = pytest.mark.gpu


def _loading_dict(L):
    out = {}
    for name, _ in L._fields_:
        v = getattr(L, name)
        out[name] = list(v) if name in ("box_c", "box_h", "hidden") else v
    return out


def _rest_inv(R, bi, bj):
    """1/L0 as the device keeps it: float(1 / |r|), |r| formed in double."""
    r = R[bj] - R[bi]
    L2 = r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1] + r[:, 2] * r[:, 2]
    return (1.0 / np.sqrt(L2)).astype(np.float32).astype(np.float64)


def oracle_of(s, R0=None):
    """The CPU restatement of system s in its current state (bond rest
    lengths from R0, default the current positions, as set_bonds took them)."""
    st = s.state()
    N = s.N
    R = st["R"].reshape(3, N).T.copy()
    b = s.bonds
    p = s.params
    return OracleNativeStep(R, s.radius.astype(np.float64), b.bi, b.bj,
                            _rest_inv(R if R0 is None else R0, b.bi, b.bj), b.br, b.sc, b.tc,
                            E=p.E, nu=p.nu, rho=p.rho, mu=p.mu, mur=p.mur, e=p.e, dt=p.dt,
                            refmode=p.refmode, loading=_loading_dict(s.loading()),
                            V=st["V"].reshape(3, N).T, W=st["W"].reshape(3, N).T)


def _flags(s):
    return s.bond_state()[1]


def _explain(o, got, ref, key, step):
    """Print what the restatement holds around the particle that deviates."""
    i = int(np.argmax(np.abs(got - ref).max(1)))
    print(f"[explain] step {step} {key}: particle {i} device {got[i]} oracle {ref[i]} "
          f"|V| {np.linalg.norm(o.V[i]):.3e} r {o.rad[i]:.4f} R {o.R[i]}", flush=True)
    for (a, b), O in o.springs.items():
        if i in (a, b):
            L = np.linalg.norm(o.R[b] - o.R[a])
            print(f"  pair {a}-{b} L {L:.6f} reach {o.rad[a] + o.rad[b]:.6f} O {O}", flush=True)
    for q in np.nonzero((o.bi == i) | (o.bj == i))[0]:
        a, b = int(o.bi[q]), int(o.bj[q])
        L = np.linalg.norm(o.R[b] - o.R[a])
        print(f"  bond {q} {a}-{b} act {o.act[q]} L {L:.6f} L0 {1 / o.iL[q]:.6f} "
              f"rsum {o.rad[a] + o.rad[b]:.6f}", flush=True)
    from scipy.spatial import cKDTree

    near = cKDTree(o.R).query_ball_point(o.R[i], 2.0 * o.rad.max())
    for j in sorted(near):
        if j != i:
            L = np.linalg.norm(o.R[j] - o.R[i])
            if L < o.rad[i] + o.rad[j] + 0.05:
                print(f"  near {j} L {L:.6f} rsum {o.rad[i] + o.rad[j]:.6f}", flush=True)


# The contact law switches from the damped spring to the Coulomb-capped force
# where |Ft| = mu |fn| (md_calculateforcesdem3.cu:295-314), and the two
# branches differ by the damping term.  A contact that sticks at the cap (its
# spring is reset to the cap each step) sits on that switch, where rounding
# picks the branch, so two implementations of the law may take different
# branches there; the comparison of a trajectory ends at the first evaluation
# in which a contact's Coulomb margin is below TIE.
TIE = 1e-9


def device_springs(ctx, s, bond_keys, K=48):
    """The contact springs of never-bonded pairs (i < j) held in the device's
    contact rows, read through the slot-layout export (type-0 slots that are
    not bond records)."""
    from rfd.reflayout import Grid, RefLayout

    N = s.N
    lay = RefLayout(ctx, np.zeros(3 * N, np.float32), K, Grid(1.0, 1.0, 1, 1, 1))
    s.export_reference_layout(lay)
    h = lay.host()
    IL = h["IL"].reshape(N, K).astype(np.int64)
    ty = h["ILtype"].reshape(N, K)
    O = h["Oijt"].reshape(N, K, 3).astype(np.float64)
    ii, kk = np.nonzero((IL != 0xFFFFFFFF) & (IL >= 0) & (ty == 0))
    jj = IL[ii, kk]
    m = (ii < jj) & ~np.isin(ii * N + jj, bond_keys) & np.any(O[ii, kk] != 0, axis=1)
    return {(int(i), int(j)): O[i, k].copy() for i, j, k in zip(ii[m], jj[m], kk[m])}


def _fref(s):
    """The bond force at unit strain for the mean radius, E pi r^2 (the force
    scale of the FIRE tolerance), and r_mean."""
    r = float(np.mean(s.radius))
    return s.params.E * np.pi * r * r, r


def run_side_by_side(s, o, steps, tol_R, tol_FM, exact_flags=True, floor=1e-12,
                     sync_bonds=False, tie=TIE):
    """Step the device system and the restatement together, each step from
    the device's positions and velocities (in double, so that the comparison
    measures one step of each and no stick-slip amplification of earlier
    last-bit differences: the Brazilian platens' friction is discontinuous at
    zero sliding speed).  The restatement keeps its own bond state, contact
    springs, broken-bond set and loading control throughout.  Compared after
    each step: positions, forces, moments, the platen reactions and every
    bond's intact/broken flag.  Ends after `steps` steps or at the first
    Coulomb tie (with sync_bonds, a step with a tie is left out and the next
    starts again from the device's state; tie is the margin below which the
    branch is decided by rounding: 1e-9 against FP64, 1e-5 against FP32
    arithmetic).  sync_bonds: the restatement also takes the device's bond
    state and flags, and the springs of its contact rows, before each step
    (the FP32 path, whose state differs from the FP64 one by its own
    rounding, which stick-slip at a Coulomb cap amplifies).  Forces must agree to tol_FM of the largest plus `floor`
    times E pi r_mean^2 (moments: times r_mean as well), the floor covering
    the rounding of stretches near zero at the start of a run.  Returns
    (breaks, most pair contacts, most flag mismatches, largest deviations,
    steps compared)."""
    fref, rm = _fref(s)
    skipped = 0
    worst = {"R": 0.0, "F": 0.0, "M": 0.0, "sums": 0.0}
    broken_total = 0
    touching = 0
    mism = 0
    done = 0
    for k in range(steps):
        st0 = s.state64()
        o.R, o.V, o.W = st0["R"], st0["V"], st0["W"]
        if sync_bonds:
            o.load_bond_state(*s.bond_state())
            o.springs = device_springs(s.ctx, s, o._bond_keys)
        s.step(1)
        s0, s1 = o.step()
        if o.cap_tie < tie:
            if not sync_bonds:
                print(f"[side by side] step {k + 1}: Coulomb margin {o.cap_tie:.1e}, comparison "
                      "ends", flush=True)
                break
            # the next step starts again from the device's whole state
            skipped += 1
            continue
        done += 1
        broken_total += len(o.broke_last)
        touching = max(touching, len(o.springs))
        st = s.state64()
        dR = np.abs(st["R"] - o.R).max()
        worst["R"] = max(worst["R"], dR)
        if dR > tol_R:
            i = int(np.argmax(np.abs(st["R"] - o.R).max(1)))
            L = s.loading()
            print(f"[explain R] step {k + 1} particle {i}: device {st['R'][i]} oracle {o.R[i]} "
                  f"V0 {st0['V'][i]} Vdev {st['V'][i]} Vorc {o.V[i]} Fdev {st['F'][i]} "
                  f"Forc {o.F[i]} top {L.top} bottom {L.bottom}", flush=True)
        assert dR <= tol_R, f"step {k + 1}: |dR| = {dR:.3e} mm"
        for key, ref, fl0 in (("F", o.F, floor * fref), ("M", o.M, floor * fref * rm)):
            got = st[key]
            scale = max(np.abs(ref).max(), 1e-300)
            dev = np.abs(got - ref).max()
            d = dev / scale
            worst[key] = max(worst[key], d)
            if dev > tol_FM * scale + fl0:
                _explain(o, got, ref, key, k + 1)
            assert dev <= tol_FM * scale + fl0, f"step {k + 1}: {key} deviates {d:.3e} of max"
        L = s.loading()
        ds = max(abs(L.sum_bottom - s0), abs(L.sum_top - s1)) / max(abs(s0), abs(s1), 1e-300)
        worst["sums"] = max(worst["sums"], ds)
        fl = _flags(s)
        diff = int(np.count_nonzero((fl == 1) != (o.act == 1)))
        mism = max(mism, diff)
        if exact_flags:
            assert diff == 0, f"step {k + 1}: {diff} bond flags differ"
    if skipped:
        print(f"[side by side] {skipped} steps with a contact within {tie:.0e} of its Coulomb "
              "switch not compared", flush=True)
    return broken_total, touching, mism, worst, done


def test_fp64_brazil_disc_steps_match_cpu_oracle(ctx):
    nc = native_cases.packed_brazil_disc(ctx, fp64=True, v_load=-20.0 / VELOCITY)
    s = nc.sys
    o = oracle_of(s)
    steps = 300
    print(f"[c1 fp64] N={s.N} bonds={s.bonds.n}", flush=True)
    nbreak, ntouch, _, worst, done = run_side_by_side(s, o, steps, tol_R=1e-12, tol_FM=1e-10)
    h = s.history()[:done]
    ho = np.array(o.history)[:done]
    assert h.shape == ho.shape == (done, 3)
    np.testing.assert_allclose(h, ho, rtol=1e-9, atol=1e-12 * np.abs(ho).max())
    L = s.loading()
    print(f"[c1 fp64] {done} steps compared: breaks {nbreak}, contact pairs {ntouch}, "
          f"worst {worst}, stress {L.stress * 1e3:.2f} MPa", flush=True)
    assert done >= 200
    assert nbreak > 0, "no bond broke: the run does not reach the break criterion"
    assert np.abs(ho[:, 1:]).max() > 0
    assert ctx.flags() == 0


def test_fp64_uniaxial_cylinder_steps_match_cpu_oracle(ctx):
    """Lattice cylinder D6 x H10 under uniaxial platens (load axis z, with
    the cut box) closing at 20 m/s, weak bonds N(2,0.4) / N(4,0.8) MPa so
    that they break at the platens within the run."""
    nc = native_cases.uniaxial_cylinder(ctx, 6.0, 10.0, seed=12, fp64=True, flat_ends=True,
                                        strength=(2e6, 4e6), v_load=-20.0 / VELOCITY)
    s = nc.sys
    s.seat_platens(preload=0.0)
    o = oracle_of(s, R0=None)
    # seat_platens moved the platens only: the rest lengths are those of the
    # positions the bonds were made at, which seat_platens leaves in place
    nbreak, ntouch, _, worst, done = run_side_by_side(s, o, 200, tol_R=1e-12, tol_FM=1e-10)
    print(f"[uniaxial fp64] N={s.N} {done} steps compared: breaks {nbreak}, contact pairs "
          f"{ntouch}, worst {worst}", flush=True)
    assert done >= 150 and nbreak > 0
    ho = np.array(o.history)[:done]
    np.testing.assert_allclose(s.history()[:done], ho, rtol=1e-9, atol=1e-12 * np.abs(ho).max())


def test_fp32_brazil_disc_tracks_cpu_oracle(ctx):
    """The FP32 path of the same disc against the FP64 restatement: forces and
    moments within 1e-4 of the largest over 300 steps (FP32 bond and contact
    arithmetic, FP64 positions), positions within 1e-6 mm, and the broken-bond
    set equal at every step (its stress criterion is rechecked in FP64 near
    the threshold)."""
    nc = native_cases.packed_brazil_disc(ctx, fp64=False, v_load=-20.0 / VELOCITY)
    s = nc.sys
    o = oracle_of(s)
    nbreak, ntouch, mism, worst, done = run_side_by_side(s, o, 300, tol_R=1e-7, tol_FM=1e-4,
                                                         floor=1e-9, sync_bonds=True,
                                                         tie=1e-5)
    print(f"[c1 fp32] N={s.N} {done} steps compared: breaks {nbreak}, contact pairs {ntouch}, "
          f"flag mismatch {mism}, worst {worst}", flush=True)
    assert done >= 200 and nbreak > 0


def test_batch_member_tracks_cpu_oracle(ctx):
    """One specimen of a batch of Brazilian discs (FP32 path, one launch over
    all discs per step) against the FP64 restatement of that disc alone, each
    step from the device's whole state (positions, velocities, bond state,
    contact springs) over 200 steps: forces within 1e-4 of the largest,
    positions within 1e-7 mm, bond flags equal."""
    case = native_cases.brazil_batch(ctx, S=3, seed=31, n_lo=1000, n_hi=3000,
                                     mu_range=(2.0, 3.0), v_load=-20.0 / VELOCITY)
    bs = case.sys
    k = 1
    a, b = int(case.first[k]), int(case.first[k + 1])
    member = native_cases.batch_member(ctx, case, k, fp64=True)
    o = oracle_of(member)
    fref, _ = _fref(member)
    member.close()
    N = bs.N
    worst = {"R": 0.0, "F": 0.0}
    nbreak = 0
    bm = (bs.bonds.bi >= a) & (bs.bonds.bi < b)
    assert np.array_equal(bs.bonds.bi[bm] - a, o.bi) and np.array_equal(bs.bonds.bj[bm] - a, o.bj)
    bkeys = np.sort(bs.bonds.bi.astype(np.int64) * N + bs.bonds.bj.astype(np.int64))
    done = 0
    for step in range(200):
        st0 = bs.state64()
        o.R, o.V, o.W = st0["R"][a:b], st0["V"][a:b], st0["W"][a:b]
        st12, fl0 = bs.bond_state()
        o.load_bond_state(st12[bm], fl0[bm])
        sp = device_springs(ctx, bs, bkeys)
        o.springs = {(i - a, j - a): O for (i, j), O in sp.items() if a <= i < b}
        bs.step(1)
        o.step()
        if o.cap_tie < 1e-5:
            continue  # the next step starts again from the device's state
        done += 1
        nbreak += len(o.broke_last)
        st = bs.state64()
        R = st["R"][a:b]
        dR = np.abs(R - o.R).max()
        worst["R"] = max(worst["R"], dR)
        assert dR <= 1e-7, f"step {step + 1}: |dR| {dR:.3e}"
        F = st["F"][a:b]
        dev = np.abs(F - o.F).max()
        d = dev / max(np.abs(o.F).max(), 1e-300)
        worst["F"] = max(worst["F"], d)
        if dev > 1e-4 * np.abs(o.F).max() + 1e-9 * fref:
            print(f"[batch] step {step + 1} Coulomb margin {o.cap_tie:.2e}", flush=True)
            _explain(o, F, o.F, "F", step + 1)
        assert dev <= 1e-4 * np.abs(o.F).max() + 1e-9 * fref, \
            f"step {step + 1}: F deviates {d:.3e} of max"
        fl = _flags(bs)[bm]
        assert np.array_equal(fl == 1, o.act == 1), f"step {step + 1}: bond flags differ"
    print(f"[batch member] N={b - a} {done} steps compared: breaks {nbreak} worst {worst}",
          flush=True)
    assert done >= 150 and nbreak > 0


def test_break_decision_at_threshold_matches_f64_oracle(ctx):
    """Bonds whose strength is their own stress rounded to FP32: the margin is
    within FP32 rounding (|margin| < 1e-7 relative), where the FP32 criterion
    alone would decide at random; the device rechecks such bonds in FP64 and
    must decide as the FP64 bond law does.  Tensile set: stretched bonds at
    rest state.  Shear set: bonds given a shear spring O (FP32), at their rest
    length.  One force evaluation, zero velocities."""
    nc = native_cases.uniaxial_cylinder(ctx, 6.0, 8.0, seed=41, strength=(20e6, 40e6))
    s = nc.sys
    N = s.N
    b = s.bonds
    R0 = s.state()["R"].reshape(3, N).T.copy()
    L0 = np.linalg.norm(R0[b.bj] - R0[b.bi], axis=1).astype(np.float32)
    g = np.random.default_rng(43)
    R1 = R0 + g.standard_normal(R0.shape) * 2e-3
    r = R1[b.bj] - R1[b.bi]
    L1 = np.sqrt(r[:, 0] * r[:, 0] + r[:, 1] * r[:, 1] + r[:, 2] * r[:, 2])
    il = (1.0 / L0.astype(np.float64)).astype(np.float32).astype(np.float64)
    E, G = s.params.E, 0.5 * s.params.E / (1 + s.params.nu)
    sig = E * (L1 - 1.0 / il) * il
    nb = b.n
    idx = np.arange(nb)
    tens = (idx % 2 == 0) & (sig > 1e-4)
    shear = idx % 2 == 1
    sc = np.full(nb, 1e3, np.float32)
    tc = np.full(nb, 1e3, np.float32)
    sc[tens] = sig[tens].astype(np.float32)
    # shear springs: a tangential O of ~1e-3 mm; the FP64 criterion sees
    # G il |O| + (tensile part only in Sc)
    n = r / L1[:, None]
    t = np.cross(n, g.standard_normal((nb, 3)))
    t /= np.linalg.norm(t, axis=1)[:, None]
    O = (t * g.uniform(5e-4, 2e-3, nb)[:, None]).astype(np.float32)
    O[~shear] = 0
    Od = O.astype(np.float64)
    Mc = G * il * np.sqrt(Od[:, 0] * Od[:, 0] + Od[:, 1] * Od[:, 1] + Od[:, 2] * Od[:, 2])
    # the device turns O by the FP32 rounding of its bond direction (about
    # 1e-8), so shear margins are drawn from +-[1e-7, 1e-6] before rounding to
    # FP32, where the FP32 criterion alone (a few 1e-7) still errs
    m = g.uniform(1e-7, 1e-6, nb) * np.where(idx % 4 == 1, 1.0, -1.0)
    tc[shear] = (Mc[shear] * (1.0 + m[shear])).astype(np.float32)
    # bonds that land within 1e-11 (tensile) or 2e-8 (shear) of the threshold
    # after rounding are left out of the check (strength 1e3)
    close_t = tens & (np.abs(sc.astype(np.float64) - sig) < 1e-11 * sig)
    close_s = shear & (np.abs(tc.astype(np.float64) - Mc) < 2e-8 * Mc)
    sc[close_t] = 1e3
    tc[close_s] = 1e3
    tens &= ~close_t
    shear &= ~close_s
    s.set_bonds(type(b)(b.bi, b.bj, b.br, sc, tc, L0))
    s.set_positions(R1)
    state = np.zeros((nb, 12), np.float32)
    state[:, 3:6] = O
    s.set_bond_state(state, np.ones(nb, np.uint8))
    s.compute_forces()
    fl = _flags(s)
    # the FP64 law on the same inputs
    Rs = np.ascontiguousarray(R1.T).reshape(-1)
    Z = np.zeros(3 * N)
    st = np.zeros(8 * nb)
    st[0::8], st[1::8], st[2::8] = Od[:, 0], Od[:, 1], Od[:, 2]
    act = np.ones(nb, np.uint8)
    F, M = np.zeros(3 * N), np.zeros(3 * N)
    brd, scd, tcd = (x.astype(np.float64) for x in (b.br, sc, tc))
    olib().orc_native_bonds_f64(ptr(Rs), ptr(Z), ptr(Z), N, ptr(b.bi), ptr(b.bj),
                                ptr(np.ascontiguousarray(il)), ptr(brd), ptr(scd), ptr(tcd),
                                ptr(st), ptr(act), nb, s.params.dt, E, G, 0, ptr(F), ptr(M))
    margin_t = (sc.astype(np.float64) - sig) / sig
    margin_s = (tc.astype(np.float64) - Mc) / Mc
    print(f"[threshold] tensile {tens.sum()} (|margin| < 1e-7: "
          f"{np.sum(np.abs(margin_t[tens]) < 1e-7)}), shear {shear.sum()} (|margin| < 3e-7: "
          f"{np.sum(np.abs(margin_s[shear]) < 3e-7)}), broken {np.sum(act == 0)}", flush=True)
    assert np.sum(act[tens] == 0) > 20 and np.sum(act[tens] == 1) > 20
    assert np.sum(act[shear] == 0) > 20 and np.sum(act[shear] == 1) > 20
    assert np.array_equal(fl == 1, act == 1), np.nonzero((fl == 1) != (act == 1))[0][:10]
    # the whole evaluation (bonds breaking now meet the contact law at once)
    p = s.params
    o = OracleNativeStep(R1, s.radius.astype(np.float64), b.bi, b.bj, il, b.br, sc, tc, E=p.E,
                         nu=p.nu, rho=p.rho, mu=p.mu, mur=p.mur, e=p.e, dt=p.dt, refmode=0)
    o.state.reshape(-1, 8)[:, 0:3] = Od
    o.forces()
    assert np.array_equal(o.act, act)
    got = s.state64()
    for key, ref in (("F", o.F), ("M", o.M)):
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got[key], ref, rtol=0, atol=1e-5 * scale, err_msg=key)
