"""Multi-step launches (rfd_system_set_persistent): the fused steps between
two contact-row rebuilds run as one cooperative launch with a grid barrier
per step and the loading control inside.  The arithmetic of each step is the
fused pass's and the control's, so the state, the bond state and flags, the
loading control and its history equal those of one launch per step, bit for
bit: over rebuilds, breaking bonds and odd step counts, for one uniaxial
cylinder, one Brazilian disc and a batch of discs whose tiles span two
specimens.  A system with more force tiles than resident workgroups runs its
steps one launch each, as before.
"""
import numpy as np
import pytest

from rfd import native_cases

from test_gpu_batch import _same_loading, _same_state
from test_gpu_fused import _brazil, _compare, _uniaxial

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["uniaxial", "brazil"])
def test_multi_step_launches_match_one_launch_per_step(ctx, mode):
    build = _uniaxial if mode == "uniaxial" else _brazil
    a, b = build(ctx), build(ctx)
    a.set_persistent(True)
    assert a.persistent() == (True, True), a.persistent()
    assert b.persistent()[0] is False
    n0 = a.bond_counts()[0]
    for n in (2, 7, 20, 33, 40, 57, 1, 64):
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


def test_multi_step_batch_matches_one_launch_per_step(ctx):
    def build():
        return native_cases.brazil_batch(ctx, S=4, seed=53, n_lo=300, n_hi=1200,
                                         v_load=-5e-3, hist_cap=4096).sys

    a, b = build(), build()
    a.set_persistent(True)
    assert a.persistent() == (True, True)
    for n in (3, 37, 80, 41):
        print(f"batch: {n} steps", flush=True)
        a.step(n)
        b.step(n)
        _same_state(a, b, label=f"batch +{n}")
        for k, (La, Lb) in enumerate(zip(a.specimen_loadings(), b.specimen_loadings())):
            _same_loading(La, Lb, f"batch specimen {k}")
            assert np.array_equal(a.specimen_history(k), b.specimen_history(k))
        sa, fa = a.bond_state()
        sb, fb = b.bond_state()
        assert np.array_equal(fa, fb) and np.array_equal(sa, sb)
    assert any(L.loading for L in a.specimen_loadings())
    assert ctx.flags() == 0
    a.close()
    b.close()


def test_multi_step_graph_capture_falls_back(ctx):
    """Graph replays capture one launch per step even with multi-step
    launches on; replays and multi-step launches interleave with the same
    bits as plain launches."""
    a, b = _brazil(ctx), _brazil(ctx)
    a.set_persistent(True)
    for n, how in ((40, "graph"), (6, "plain"), (34, "plain"), (80, "graph")):
        print(f"{how}: {n} steps", flush=True)
        if how == "graph":
            a.step_graph(n)
        else:
            a.step(n)
        b.step(n)
        _compare(a, b, f"{how} +{n}")
    a.close()
    b.close()


def test_large_system_does_not_qualify(ctx):
    """A cylinder with more force tiles than the GPU holds resident
    workgroups reports that multi-step launches do not apply, and steps as
    before."""
    a = native_cases.uniaxial_cylinder(ctx, 56.0, 60.0, seed=3).sys
    b = native_cases.uniaxial_cylinder(ctx, 56.0, 60.0, seed=3).sys
    a.set_persistent(True)
    a.step(2)
    b.step(2)
    on, applies = a.persistent()
    print(f"N = {a.N}: requested {on}, applies {applies}", flush=True)
    assert on and not applies
    a.step(4)
    b.step(4)
    _compare(a, b, "large")
    a.close()
    b.close()
