"""GPU parity of the superseded per-slot force path and penalty platens
(rfd_forces_slots: md_calculateforcesdem.cu d_CalculateForcesDEM :328,
_1 :501, _2 :581; rfd_platen_uniaxial_penalty: md_uniaxialcompression_simple.cu
:29) against the oracle, bit for bit (the platen reaction sums, FP64 sums of
the same FP32 terms in another order, to 1e-12 relative):

* each mode once, on a jittered lattice cylinder with random velocities and
  spins, where a third of the bonded slots are turned into contact slots (so
  that about half of those overlap and feel the Hertz-Mindlin law);
* 30 steps of the per-slot step: bond slots, contact slots, sum-up, the
  stand-alone break check with strengths low enough that bonds break,
  penalty platens pushing on both end layers, increments.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleLayout
from rfd import cases
from rfd.reflayout import RefLayout

from test_gpu_reflayout import assert_same

pytestmark = pytest.mark.gpu

SLOT_KEYS = ("R", "V", "W", "F", "M", "ILtype", "rij", "Oijt", "Fijn", "Fijt", "Mijn", "Mijt",
             "Mijadd")


def _twin(ctx, seed=2):
    case = cases.reference_uniaxial(6.0, 8.0, seed=seed)
    m = case.mat
    aacut = float(np.float32(m.a_farcut) ** 2)
    g, o = RefLayout(ctx, case.R, case.K, case.grid()), OracleLayout(case.R, case.K, case.grid())
    rng = np.random.default_rng(seed)
    V = (rng.standard_normal(3 * case.N) * 0.05).astype(np.float32)
    W = (rng.standard_normal(3 * case.N) * 0.2).astype(np.float32)
    for s in (g, o):
        s.cell_distribution()
        s.il_construct(aacut)
        s.bonds_construct(m.b_dd)
    ILtype = o.ILtype.copy()
    bonded = np.flatnonzero(ILtype == 1)
    ILtype[rng.choice(bonded, len(bonded) // 3, replace=False)] = 0
    o.ILtype[:] = ILtype
    o.V[:] = V
    o.W[:] = W
    g.ILtype.copy_(torch.from_numpy(ILtype).cuda())
    g.V.copy_(torch.from_numpy(V).cuda())
    g.W.copy_(torch.from_numpy(W).cuda())
    return case, g, o


def test_each_mode_bitexact(ctx):
    case, g, o = _twin(ctx)
    m = case.mat
    for mode in (2, 1, 0):
        for s in (g, o):
            s.forces_slots(mode, m)
            s.sumup_forces()
        assert_same(g.host(), o.host(), keys=SLOT_KEYS, where=f"mode {mode}")
    # the contact law acted: overlapping contact slots carry a normal force
    h = o.host()
    contact = np.repeat((h["ILtype"] == 0) & (h["IL"] != 0xFFFFFFFF), 3)
    assert np.count_nonzero(h["Fijn"][contact]) > 100
    assert ctx.flags() == 0


def test_per_slot_steps_bitexact(ctx):
    case, g, o = _twin(ctx, seed=6)
    m = case.mat
    cmp = case.cmp
    c = cmp.center_d
    z = case.R.reshape(3, -1)[2]
    # the end layers start 0.1 mm inside the platens
    Hd2 = float(min(z.max() - c[2], c[2] - z.min())) - 0.1
    stiffness = m.E * m.p_r
    sc, tc = 0.02 * m.E, 0.04 * m.E
    n0 = int((o.ILtype == 1).sum())
    for step in range(30):
        sums = []
        for s in (g, o):
            s.forces_slots(2, m)
            s.forces_slots(1, m)
            s.sumup_forces()
            s.check_break(sc, tc, m.b_r)
            sums.append(s.platen_uniaxial_penalty(c, cmp.RR, Hd2, stiffness, m.p_mu, 1e30))
            s.increments(m)
        np.testing.assert_allclose(sums[0], sums[1], rtol=1e-12)
        if step % 10 == 9:
            assert_same(g.host(), o.host(), keys=SLOT_KEYS, where=f"step {step}")
    assert_same(g.host(), o.host(), keys=SLOT_KEYS, where="end")
    n1 = int((o.ILtype == 1).sum())
    print(f"bond slots {n0} -> {n1}; last reactions {sums[1]}", flush=True)
    assert n1 < n0
    assert sums[1][0] < 0 < sums[1][1]  # (bottom, top): the platens push inwards
    assert ctx.flags() == 0
