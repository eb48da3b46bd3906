"""The reference slot layout in and out of the native tier
(rfd_system_import_reference_layout / rfd_system_export_reference_layout).

Source: a reference-layout (tier-1) uniaxial run under fast loading, so that
bonds carry accumulated springs and moments.

* Import gives the same bonds, bond state, positions and velocities as the
  host-side conversion (bonds_from_slots + set_bonds + set_bond_state), bit
  for bit.
* Export after import: positions and velocities come back bit for bit; every
  row holds the same bonded partners as the tier-1 row; the owner slot
  (i < j) holds tier-1's Oijt and Mijt bit for bit, 1/iL within one ulp (the
  native tier keeps the rest length 1/iL and takes its reciprocal), Mijn as the
  twist along the bond (tier-1 keeps a vector within 1e-3 of that line) and
  rij as the direction at R - V dt (within 1e-3); the other end's slot holds
  the negated state, which is what tier-1's own second slot holds.
* Contact springs survive: every tier-1 type-0 slot with a spring comes back
  with the same spring bits.
* Rows longer than K are an argument error.
"""
import numpy as np
import pytest

from rfd import cases, native
from rfd.reflayout import RefLayout, UniaxialDriver

pytestmark = pytest.mark.gpu


def _tier1(ctx, steps=300):
    c = cases.reference_uniaxial(8.0, 14.0, seed=21)
    c.cmp.v_load *= 150
    d = UniaxialDriver(ctx, c.R, c.mat, c.box, c.cmp, c.K)
    d.setup()
    for _ in range(steps):
        d.step()
    return c, d


def _rows(h, N, K, ty):
    IL = h["IL"].reshape(N, K)
    t = h["ILtype"].reshape(N, K)
    out = {}
    for i, k in zip(*np.nonzero((IL != 0xFFFFFFFF) & (t == ty))):
        out[(int(i), int(IL[i, k]))] = int(i) * K + int(k)
    return out


def test_import_and_export_reference_layout(ctx):
    c, d = _tier1(ctx)
    m, N, K = c.mat, c.N, c.K
    h = d.s.host()
    R64 = h["R"].astype(np.float64).reshape(3, N).T
    p = native.params_from_material(m, refmode=1)
    a = native.NativeSystem(ctx, p, R64, np.full(N, m.p_r, np.float32))
    a.import_reference_layout(d.s, m)
    b = native.NativeSystem(ctx, p, R64, np.full(N, m.p_r, np.float32))
    bb, st, fl = native.bonds_from_slots(h, N, K, m)
    b.set_bonds(bb)
    b.set_bond_state(st, fl)
    b.set_velocities(h["V"], h["W"])
    sa, fa = a.bond_state()
    sb, fb = b.bond_state()
    print(f"N={N} bonds={len(fa)}", flush=True)
    assert len(fa) > 0 and np.array_equal(fa, fb) and np.array_equal(sa, sb)
    xa, xb = a.state(), b.state()
    for k in ("R", "V", "W"):
        assert np.array_equal(xa[k], xb[k]), k

    lay = RefLayout(ctx, h["R"], K, d.grid)
    a.export_reference_layout(lay)
    e = lay.host()
    for k in ("R", "V", "W"):
        assert np.array_equal(e[k], h[k]), k
    b1, e1 = _rows(h, N, K, 1), _rows(e, N, K, 1)
    assert set(b1) == set(e1)
    v = lambda arr, sl: arr.reshape(-1, 3)[sl]
    own = [(i, j) for (i, j) in b1 if i < j]
    si = np.array([b1[q] for q in own])
    se = np.array([e1[q] for q in own])
    # the native tier keeps the rest length 1/iL and its own reciprocal
    np.testing.assert_array_max_ulp(h["iL"][si], e["iL"][se], maxulp=1)
    for k in ("Oijt", "Mijt"):
        assert np.array_equal(v(h[k], si), v(e[k], se)), k
    mn_h, mn_e = v(h["Mijn"], si), v(e["Mijn"], se)
    np.testing.assert_allclose(mn_e, mn_h, rtol=0, atol=1e-3 * np.abs(mn_h).max() + 1e-30)
    np.testing.assert_allclose(v(e["rij"], se), v(h["rij"], si), rtol=0, atol=1e-3)
    # the other end's slot: the negated state, as tier-1 keeps it
    sm_h = np.array([b1[(j, i)] for (i, j) in own])
    sm_e = np.array([e1[(j, i)] for (i, j) in own])
    for k in ("Oijt", "Mijt", "Mijn", "rij"):
        assert np.array_equal(v(e[k], sm_e), -v(e[k], se)), k
        np.testing.assert_allclose(v(h[k], sm_h), -v(h[k], si), rtol=0,
                                   atol=1e-6 * (np.abs(v(h[k], si)).max() + 1e-30), err_msg=k)
    # contact springs
    c0, e0 = _rows(h, N, K, 0), _rows(e, N, K, 0)
    spring = [q for q, sl in c0.items() if np.any(v(h["Oijt"], sl) != 0)]
    print(f"contact slots with a spring: {len(spring)}", flush=True)
    for q in spring:
        assert q in e0, q
        assert np.array_equal(v(h["Oijt"], c0[q]), v(e["Oijt"], e0[q])), q
    # too few slots per row
    small = RefLayout(ctx, h["R"], 2, d.grid)
    with pytest.raises(native.RfdError):
        a.export_reference_layout(small)
    assert ctx.flags() == 0
    a.close()
    b.close()
