"""Explicit steps with the particle update fused into the force pass.

The fused pass runs the separate particle pass's arithmetic on the summed
loads in registers, and the platen reactions are fixed-point sums, so both
give the same bits: positions, velocities, spins, loads, bond state and
flags, the loading control and its history.  Checked over steps with
contact-row rebuilds, breaking bonds, odd step counts (the last step of an
odd count runs the separate pass) and graph replays, for both loading modes
and for batched specimens.
"""
import numpy as np
import pytest

from rfd import native_cases

from test_gpu_batch import _same_loading, _same_state

pytestmark = pytest.mark.gpu


def _uniaxial(ctx):
    # weak bonds and a fast platen, so that bonds break within the test
    s = native_cases.uniaxial_cylinder(ctx, 8.0, 16.0, seed=9, strength=(4e6, 8e6),
                                       v_load=-0.05).sys
    s.seat_platens(0.05)
    return s


def _brazil(ctx):
    case = native_cases.brazil_batch(ctx, S=1, seed=13, n_lo=1500, n_hi=1600,
                                     v_load=-5e-3, hist_cap=4096)
    return native_cases.batch_member(ctx, case, 0)


def _compare(a, b, label):
    _same_state(a, b, label=label)
    _same_loading(a.loading(), b.loading(), label)
    sa, fa = a.bond_state()
    sb, fb = b.bond_state()
    assert np.array_equal(fa, fb), f"{label}: bond flags differ"
    assert np.array_equal(sa, sb), f"{label}: bond state differs"
    assert np.array_equal(a.history(), b.history()), f"{label}: history differs"


@pytest.mark.parametrize("mode", ["uniaxial", "brazil"])
def test_fused_steps_match_separate_pass(ctx, mode):
    build = _uniaxial if mode == "uniaxial" else _brazil
    a, b = build(ctx), build(ctx)
    b.set_fusion(False)
    assert a.fused() and not b.fused()
    n0 = a.bond_counts()[0]
    for n in (1, 2, 7, 20, 33, 40, 57):
        print(f"{mode}: {n} steps", flush=True)
        a.step(n)
        b.step(n)
        _compare(a, b, f"{mode} after +{n}")
    assert a.bond_counts() == b.bond_counts()
    if mode == "uniaxial":
        assert a.loading().loading and a.bond_counts()[0] < n0
    assert ctx.flags() == 0
    a.close()
    b.close()


def test_fused_graph_replays_match_separate_pass(ctx):
    a, b = _uniaxial(ctx), _uniaxial(ctx)
    b.set_fusion(False)
    for n, how in ((40, "graph"), (3, "plain"), (37, "graph"), (80, "graph")):
        print(f"{how}: {n} steps", flush=True)
        if how == "graph":
            a.step_graph(n)
        else:
            a.step(n)
        b.step(n)
        _compare(a, b, f"{how} +{n}")
    assert ctx.flags() == 0
    a.close()
    b.close()


def test_fused_batch_matches_separate_pass(ctx):
    """Batched discs: tiles that span two specimens split their platen sums
    between them; the fused batch equals the separate-pass batch bit for bit,
    per-specimen loading and history included."""

    def build():
        return native_cases.brazil_batch(ctx, S=4, seed=53, n_lo=300, n_hi=1200,
                                         v_load=-5e-3, hist_cap=4096).sys

    a, b = build(), build()
    b.set_fusion(False)
    assert a.fused() and not b.fused()
    for n, how in ((3, "plain"), (37, "plain"), (80, "graph"), (41, "plain")):
        print(f"batch {how}: {n} steps", flush=True)
        if how == "graph":
            a.step_graph(n)
        else:
            a.step(n)
        b.step(n)
        _same_state(a, b, label=f"batch {how} +{n}")
        for k, (La, Lb) in enumerate(zip(a.specimen_loadings(), b.specimen_loadings())):
            _same_loading(La, Lb, f"batch specimen {k}")
            assert np.array_equal(a.specimen_history(k), b.specimen_history(k))
        assert np.array_equal(a.bond_state()[1], b.bond_state()[1])
    assert any(L.loading for L in a.specimen_loadings())
    assert ctx.flags() == 0
    a.close()
    b.close()
