"""The full-size batch: 256 Brazilian discs (particle counts log-uniform in
[1k, 40k], about 2.7 M particles and 15 M bonds), stepped 400 explicit steps
(20 contact-row rebuilds) as one system, then checked by properties that do
not depend on its size:

* no device condition raised (cell range, full contact rows); every bond is
  intact, broken or removed;
* each disc's force pass balances: the bond and contact loads of a disc sum
  to zero (Newton's third law) within 1e-5 of the sum of their magnitudes;
* sampled discs (the smallest, the largest, one in the middle whose first
  particle does not start a force tile) stepped alone, with the batch's cell
  grid, equal their slice of the batch bit for bit: positions, velocities,
  spins, loads, loading control and stress history.
"""
import numpy as np
import pytest

from rfd import native_cases

from test_gpu_batch import _same_loading, _same_state

pytestmark = pytest.mark.gpu

STEPS = 400


def test_full_batch_properties_and_members(ctx):
    case = native_cases.brazil_batch(ctx, S=256)
    s = case.sys
    S = len(case.loadings)
    n = np.diff(case.first)
    print(f"batch: {S} discs, N = {s.N}, bonds = {s.bonds.n}, discs of {n.min()}..{n.max()}",
          flush=True)
    mid = next(k for k in range(S // 2, S) if case.first[k] % 64 != 0)
    picks = sorted({int(np.argmin(n)), int(np.argmax(n)), mid})
    alone = {k: native_cases.batch_member(ctx, case, k) for k in picks}
    s.step(STEPS)
    Ls = s.specimen_loadings()
    assert ctx.flags() == 0
    intact, broken, removed = s.bond_counts()
    assert intact + broken + removed == s.bonds.n
    print(f"after {STEPS} steps: bonds intact {intact}, broken {broken}, removed {removed}; "
          f"{sum(L.loading for L in Ls)} discs loading", flush=True)
    for k, a in alone.items():
        a.step(STEPS)
        lo, hi = int(case.first[k]), int(case.first[k + 1])
        print(f"disc {k}: N = {hi - lo}, first particle {lo}", flush=True)
        _same_state(s, a, slice(lo, hi), slice(None), f"disc {k}")
        _same_loading(Ls[k], a.loading(), f"disc {k}")
        hb, ha = s.specimen_history(k), a.history()
        assert hb.shape == (STEPS, 3)
        assert np.array_equal(hb, ha), f"disc {k}: history differs"
        assert Ls[k].stress > 0.0
        a.close()
    # the force pass alone (no platens): each disc's loads sum to zero
    s.compute_forces()
    F = s.state()["F"].reshape(3, s.N).astype(np.float64)
    worst = 0.0
    for k in range(S):
        lo, hi = int(case.first[k]), int(case.first[k + 1])
        f = F[:, lo:hi]
        net = np.abs(f.sum(axis=1)).max()
        scale = np.abs(f).sum()
        assert scale > 0.0, f"disc {k} carries no load"
        worst = max(worst, net / scale)
        assert net <= 1e-5 * scale, f"disc {k}: net force {net:.3e} of {scale:.3e}"
    print(f"force balance: worst disc {worst:.2e} of its summed |F|", flush=True)
    assert ctx.flags() == 0
    s.close()
