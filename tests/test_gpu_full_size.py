"""The headline workload at its full size: the D100 x H200 mm uniaxial cylinder
(2,159,925 particles, r ~ U[0.4, 0.6] mm, 12,793,893 bonds, strengths
N(20, 4) / N(40, 8) MPa), through properties that do not depend on size.

* The FP32 bond law against the CPU FP64 oracle (orc_native_bonds_f64 through
  the step restatement) for every bond of the specimen: two evaluations at
  fixed positions with large random velocities, the bond flags equal after
  each, forces and moments within 1e-4 of the largest component.
* Newton's third law over the whole specimen: the loads of one evaluation sum
  to zero (relative to the summed magnitudes: 1e-6, the rounding of FP32
  per-particle sums).
* The explicit step is the same bits for force tiles of 64 and 32 particles,
  and with the particle update fused into the force pass or run as a separate
  pass, over 40 steps with two contact-row rebuilds.
* The reference-layout tier on the reference packing of the same cylinder
  (2,159,925 particles on a jittered lattice, 32 slots per particle): setup
  lists, then 12 whole explicit steps from random velocities that break about
  a sixth of the bonds, bit for bit against the CPU oracle (its OpenMP build)
  in every slot and particle array, platen sums within 1e-12.
"""
import numpy as np
import pytest

from rfd import native_cases

pytestmark = pytest.mark.gpu

D, H = 100.0, 200.0


def _build(ctx):
    print(f"building the D{D:g}xH{H:g} cylinder", flush=True)
    return native_cases.uniaxial_cylinder(ctx, D, H).sys


def test_full_size_bond_law_matches_f64_oracle(ctx):
    from test_gpu_native_oracle import oracle_of

    s = _build(ctx)
    N = s.N
    assert N == 2159925 and s.bonds.n == 12793893
    g = np.random.default_rng(21)
    V = (g.standard_normal(3 * N) * 5e-3).astype(np.float32)
    W = (g.standard_normal(3 * N) * 1e-2).astype(np.float32)
    s.set_velocities(V, W)
    o = oracle_of(s)
    for it in range(2):
        s.compute_forces()
        print(f"evaluation {it}: device done, oracle ...", flush=True)
        o.forces()
        fl = s.bond_state()[1]
        assert np.array_equal(fl == 1, o.act == 1), f"bond flags differ at evaluation {it}"
        out = s.state64()
        for k, ref in (("F", o.F), ("M", o.M)):
            scale = np.abs(ref).max()
            np.testing.assert_allclose(out[k], ref, rtol=0, atol=1e-4 * scale,
                                       err_msg=f"{k} at evaluation {it}")
        print(f"evaluation {it}: {int(np.count_nonzero(o.act == 0))} bonds broken", flush=True)
    F = s.state()["F"].reshape(3, N)
    np.testing.assert_allclose(F.sum(1), 0.0, atol=1e-6 * np.abs(F).sum())
    assert ctx.flags() == 0
    s.close()


def test_full_size_steps_same_bits_across_tiles_and_fusion(ctx):
    from test_gpu_batch import _same_state

    a = _build(ctx)
    b = _build(ctx)
    b.set_tiles(32, 0, True)
    c = _build(ctx)
    c.set_fusion(False)
    assert a.fused() and b.fused() and not c.fused()
    for k in (20, 20):
        for s in (a, b, c):
            s.step(k)
        print(f"{k} steps on each", flush=True)
        _same_state(a, b, label="tile 64 vs 32")
        _same_state(a, c, label="fused vs separate pass")
    fa = a.bond_state()[1]
    assert np.array_equal(fa, b.bond_state()[1]) and np.array_equal(fa, c.bond_state()[1])
    assert a.loading().steps == c.loading().steps == 40
    assert ctx.flags() == 0
    for s in (a, b, c):
        s.close()


def test_full_size_reference_layout_steps_bitexact(ctx):
    import torch

    from oracle.oracle import OracleUniaxialDriver
    from rfd import cases
    from rfd.reflayout import UniaxialDriver
    from test_gpu_reflayout import assert_same, perturb

    cg, co = cases.reference_uniaxial(D, H, seed=2), cases.reference_uniaxial(D, H, seed=2)
    N = cg.N
    assert N == 2159925 and cg.K == 32
    for c in (cg, co):
        c.cmp.v_load *= 50
    dg = UniaxialDriver(ctx, cg.R, cg.mat, cg.box, cg.cmp, cg.K)
    do = OracleUniaxialDriver(co.R, co.mat, co.box, co.cmp, co.grid(), co.K, omp=True)
    dg.setup()
    do.setup()
    print("setup done on both sides", flush=True)
    assert_same(dg.s.host(), do.s.host(), where="setup")
    b0 = int((do.s.ILtype == 1).sum())
    dv, dw = perturb(N, 51, 1e-2), perturb(N, 52, 1e-2)
    dg.s.V.copy_(torch.from_numpy(dv).cuda())
    dg.s.W.copy_(torch.from_numpy(dw).cuda())
    do.s.V[:] = dv
    do.s.W[:] = dw
    for k in range(12):
        np.testing.assert_allclose(dg.step(), do.step(), rtol=1e-12, atol=1e-300)
        if k % 6 == 5:
            assert_same(dg.s.host(), do.s.host(), where=f"step {k}")
            print(f"step {k}: bond slots {b0} -> {int((do.s.ILtype == 1).sum())}", flush=True)
    b1 = int((do.s.ILtype == 1).sum())
    assert 0.7 * b0 < b1 < 0.95 * b0
    assert ctx.flags() == 0 and do.s.last_status == 0
    assert cg.cmp.stress == pytest.approx(co.cmp.stress, rel=1e-12)
