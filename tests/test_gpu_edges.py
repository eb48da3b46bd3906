"""Edge cases of both tiers on the GPU.

* Tier 1 refuses empty inputs (N = 0) with an argument error and a message,
  and the context stays usable.
* The native tier rejects a system of no particles, and steps a system with
  no bonds at all (a packed granular sample, contacts only) and one whose
  every bond is broken (contact law through the bond records): the fused and
  separate particle updates give the same bits, and the loads of the whole
  sample sum to zero (Newton's third law; relative to the summed magnitudes:
  1e-6, the rounding of FP32 per-particle sums).
* Particle counts that do not fill the last force tile (N mod 64 != 0).
"""
import numpy as np
import pytest
import torch

from rfd import native, native_cases
from rfd._lib import RfdError

pytestmark = pytest.mark.gpu


def test_tier1_rejects_empty_inputs(ctx):
    f = torch.zeros(3, device="cuda")
    for name, args in (("rfd_zero_fm", (f.data_ptr(), f.data_ptr(), 0)),
                       ("rfd_increments", (f.data_ptr(),) * 5 + (0,) + (1.0,) * 3)):
        with pytest.raises(RfdError) as e:
            ctx.call(name, *args)
        assert e.value.status < 0 and "bad arguments" in str(e.value), name
    # the context still works
    ctx.call("rfd_zero_fm", f.data_ptr(), f.data_ptr(), 1)
    ctx.sync()
    assert ctx.flags() == 0


def test_native_rejects_no_particles(ctx):
    m = native_cases.material_c(0.4)
    p = native.params_from_material(m, refmode=0)
    with pytest.raises(RfdError):
        native.NativeSystem(ctx, p, np.zeros((0, 3)), np.zeros(0, np.float32))


def _granular(ctx, seed, bonds_broken=False):
    pk = native_cases.packed_cylinder(ctx, 8.0, 8.0, 0.6, seed=seed)
    m = native_cases.material_c(0.4)
    p = native.params_from_material(m, refmode=0, contact_capacity=32)
    s = native.NativeSystem(ctx, p, pk.R, pk.radius.astype(np.float32))
    if bonds_broken:
        pairs = s.find_pairs(0.05 * 0.5)
        n = len(pairs)
        sc, tc = native.strengths(np.random.default_rng(seed), n)
        s.set_bonds(native.Bonds(pairs[:, 0].copy(), pairs[:, 1].copy(),
                                 np.full(n, 0.4, np.float32), sc, tc))
        st, fl = s.bond_state()
        s.set_bond_state(st, np.zeros_like(fl))
    else:
        s.set_bonds(native.Bonds(*(np.zeros(0, t) for t in (np.uint32, np.uint32, np.float32,
                                                             np.float32, np.float32))))
    g = np.random.default_rng(seed)
    s.set_velocities((g.standard_normal(3 * s.N) * 2e-3).astype(np.float32), None)
    return s


@pytest.mark.parametrize("broken", [False, True])
def test_contacts_only(ctx, broken):
    a, b = _granular(ctx, 17, broken), _granular(ctx, 17, broken)
    b.set_fusion(False)
    print(f"N={a.N} (N mod 64 = {a.N % 64}), bonds={a.bonds.n}, fused={a.fused()}", flush=True)
    assert a.fused() and not b.fused()
    assert a.N % 64 != 0
    for n in (1, 30, 31):
        a.step(n)
        b.step(n)
        sa, sb = a.state(), b.state()
        for k in ("R", "V", "W", "F", "M"):
            assert np.array_equal(sa[k], sb[k]), (n, k)
    F = sa["F"].reshape(3, -1)
    assert np.abs(F).sum() > 0  # contacts carry load
    np.testing.assert_allclose(F.sum(1), 0.0, atol=1e-6 * np.abs(F).sum())
    if broken:
        assert a.bond_counts()[0] == 0
    assert ctx.flags() == 0
    a.close()
    b.close()


@pytest.mark.parametrize("fp64", [False, True], ids=["fp32", "fp64"])
@pytest.mark.parametrize("n", [2, 3, 70])
def test_one_launch_steps_of_tiny_systems(ctx, n, fp64):
    """Chains of 2, 3 and 70 bonded spheres (one partial force tile, two
    tiles) between uniaxial platens that close on them: the one-launch steps
    of small systems (the particle update on the force pass's sums, the
    loading control by the pass's last workgroup) against the same system with
    the separate particle pass and control launch.  The state is the same
    bits; the loading history agrees to 1e-12 of its largest stress (FP64
    sums its platen reactions per workgroup of the launch that forms them)."""
    from test_gpu_fire_oracle import _chain

    a, _ = _chain(ctx, n, fp64)
    b, _ = _chain(ctx, n, fp64)
    for s in (a, b):
        L = s.loading()
        L.v_load = -2e-3
        L.v = L.v_load
        s.set_loading(L)
    b.set_fusion(False)
    assert a.fused() and not b.fused()
    for k in (2, 40, 38):
        a.step(k)
        for _ in range(k):
            b.step(1)
        sa, sb = a.state(), b.state()
        for key in ("R", "V", "W", "F", "M"):
            assert np.array_equal(sa[key], sb[key]), (n, fp64, k, key)
    ha, hb = a.history(), b.history()
    assert ha.shape == hb.shape == (80, 3)
    assert np.abs(hb[:, 1:]).max() > 0  # the platens carry load
    np.testing.assert_allclose(ha, hb, rtol=0, atol=1e-12 * np.abs(hb).max())
    assert ctx.flags() == 0
    a.close()
    b.close()
