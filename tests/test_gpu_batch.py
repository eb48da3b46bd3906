"""Batched specimens, step graphs and the launch-shape autotuner on the GPU.

* A batch of Brazilian discs steps every specimen exactly as that specimen
  stepped alone (same place, same cell grid, same bonds and loading control):
  positions, velocities, forces, moments, the loading control and the
  stress-strain history are equal bit for bit.  The single-specimen path is
  itself pinned to the oracles by the other native-tier tests.
* Graph replays equal plain launches bit for bit; the host reads the state
  back after every call (eager work and synchronisation between replays), and
  a change of loading forces a new capture mid-run.
* Every force-tile / row-block shape gives the same trajectory bit for bit,
  autotuning leaves the trajectory unchanged, and the choice is cached.
"""
import numpy as np
import pytest

from rfd import native_cases

pytestmark = pytest.mark.gpu

_LOAD_FIELDS = ("stress", "strain", "fracture_stress", "fracture_strain", "top", "top_r",
                "size", "v", "sum_bottom", "sum_top", "loading", "fracture", "steps",
                "hist_len")


def _same_state(a, b, sl_a=slice(None), sl_b=slice(None), label=""):
    Na, Nb = a.N, b.N
    sa, sb = a.state(), b.state()
    for k in ("R", "V", "W", "F", "M"):
        xa = sa[k].reshape(3, Na)[:, sl_a]
        xb = sb[k].reshape(3, Nb)[:, sl_b]
        assert np.array_equal(xa, xb), f"{label}: {k} differs (max {np.abs(xa - xb).max()})"


def _same_loading(La, Lb, label=""):
    for f in _LOAD_FIELDS:
        assert getattr(La, f) == getattr(Lb, f), f"{label}: loading.{f} {getattr(La, f)} " \
                                                 f"!= {getattr(Lb, f)}"


def _small_batch(ctx, S=3, seed=31, v_load=-2.0e-3):
    return native_cases.brazil_batch(ctx, S=S, seed=seed, n_lo=400, n_hi=1500,
                                     v_load=v_load, hist_cap=4096)


def test_batch_matches_specimens_alone(ctx):
    case = _small_batch(ctx)
    S = len(case.loadings)
    alone = [native_cases.batch_member(ctx, case, k) for k in range(S)]
    steps = 130
    print(f"batch of {S}: N = {np.diff(case.first).tolist()}, {steps} steps", flush=True)
    case.sys.step(steps)
    Ls = case.sys.specimen_loadings()
    for k, s in enumerate(alone):
        s.step(steps)
        a, b = int(case.first[k]), int(case.first[k + 1])
        _same_state(case.sys, s, slice(a, b), slice(None), f"specimen {k}")
        _same_loading(Ls[k], s.loading(), f"specimen {k}")
        hb, hs = case.sys.specimen_history(k), s.history()
        assert hb.shape == (steps, 3)
        assert np.array_equal(hb, hs), f"specimen {k}: history differs"
        # the platens did work on every disc
        assert Ls[k].stress > 0.0
    for s in alone:
        s.close()
    case.sys.close()


def _pair(build):
    return build(), build()


def _graph_vs_eager(a, b, plan, loading_change=None):
    """a runs plain launches, b graph replays; compare after every call."""
    for n, label in plan:
        a.step(n)
        b.step_graph(n)
        _same_state(a, b, label=label)
        if a.S:
            for k, (La, Lb) in enumerate(zip(a.specimen_loadings(), b.specimen_loadings())):
                _same_loading(La, Lb, f"{label} specimen {k}")
        else:
            _same_loading(a.loading(), b.loading(), label)
        if loading_change is not None and label == "change":
            loading_change(a)
            loading_change(b)


def test_graph_replays_match_plain_steps_single(ctx):
    def build():
        c = native_cases.uniaxial_cylinder(ctx, 10.0, 20.0, seed=5, rebuild_every=20)
        return c.sys

    a, b = _pair(build)
    assert a.params.rebuild_every == 20

    def change(s):
        L = s.loading()
        L.v_load *= 2.0
        s.set_loading(L)

    # steps 0-39 plain; 40-119 two replays, 120-134 plain (one rebuild: the
    # row buffers swap, so 140 needs a capture of its own); 140-219 replays;
    # the loading change at 220 forces another capture
    plan = [(7, "eager start"), (33, "to the first rebuild"), (95, "two replays + 15"),
            (5, "eager"), (40, "replay 3"), (40, "replay 4"),
            (40, "change"), (80, "recaptured"), (40, "replay after recapture")]
    _graph_vs_eager(a, b, plan, change)
    a.close()
    b.close()


def test_graph_replays_match_plain_steps_batch(ctx):
    a, b = _pair(lambda: _small_batch(ctx, S=3, seed=47).sys)

    def change(s):
        Ls = s.specimen_loadings()
        for L in Ls:
            L.v_load *= 1.5
        s.set_specimens(s.spec_first, Ls, 4096)

    plan = [(1, "first"), (39, "aligned"), (80, "two replays"), (40, "change"),
            (120, "recaptured"), (13, "tail")]
    _graph_vs_eager(a, b, plan, change)
    a.close()
    b.close()


def test_tile_shapes_and_autotune_leave_results_unchanged(ctx):
    def build():
        return native_cases.uniaxial_cylinder(ctx, 12.0, 24.0, seed=9, rebuild_every=20).sys

    ref = build()
    ref.step(75)
    for fp, nb, once in ((32, 64, True), (128, 256, True), (64, 128, False), (64, 128, True),
                         (64, 0, True)):
        s = build()
        s.set_tiles(fp, nb, once)
        s.step(75)
        _same_state(ref, s, label=f"tile {fp} block {nb} single {once}")
        _same_loading(ref.loading(), s.loading(), f"tile {fp} block {nb}")
        s.close()
    # autotune mid-run: trial passes must not move the trajectory
    t = build()
    t.step(30)
    r = t.autotune(use_cache=False)
    d = r.as_dict()
    print("autotune", d, flush=True)
    assert r.force_tile in (32, 64, 128) and r.nbr_block in (0, 64, 128, 256)
    assert not r.cached
    assert all(v > 0 for v in d["force_ms"].values())
    assert all(v > 0 for v in d["nbr_ms"].values())
    key = str(r.force_tile) if r.single_eval else "64 every half"
    assert d["force_ms"][key] == min(d["force_ms"].values())
    t.step(45)
    _same_state(ref, t, label="after autotune")
    _same_loading(ref.loading(), t.loading(), "after autotune")
    # the same size bucket on the same device comes from the table
    u = build()
    r2 = u.autotune(use_cache=True)
    assert r2.cached and (r2.force_tile, r2.nbr_block, r2.single_eval) == \
        (r.force_tile, r.nbr_block, r.single_eval)
    for s in (ref, t, u):
        s.close()


def test_single_evaluation_matches_every_half_with_breaking_bonds(ctx):
    """Weak bonds under large random velocities break during the run, so the
    broken-bond contact law and the break bookkeeping run in both passes:
    the single-evaluation pass, the every-half pass and every tile shape give
    the same trajectory bit for bit."""
    def build():
        nc = native_cases.uniaxial_cylinder(ctx, 8.0, 12.0, seed=6, strength=(2e6, 4e6))
        g = np.random.default_rng(2)
        nc.sys.set_velocities((g.standard_normal(3 * nc.sys.N) * 5e-2).astype(np.float32),
                              (g.standard_normal(3 * nc.sys.N) * 5e-2).astype(np.float32))
        return nc.sys

    ref = build()
    ref.set_tiles(64, 128, False)
    ref.step(60)
    counts = ref.bond_counts()
    assert counts[1] > 0, counts  # bonds broke
    for fp, once, nb in ((64, True, 128), (32, True, 128), (128, True, 0), (32, False, 0)):
        s = build()
        s.set_tiles(fp, nb, once)
        s.step(60)
        _same_state(ref, s, label=f"tile {fp} single {once}")
        assert s.bond_counts() == counts
        s.close()
    ref.close()


@pytest.mark.parametrize("skin", [0.05, 0.35])
def test_row_search_by_cell_blocks_matches_per_particle(ctx, skin):
    """The contact rows of the cell-block search (LDS-staged candidates; with a
    0.35 mm skin the cells hold about 9 particles and many blocks exceed the
    staged copy and read candidates from global memory) equal those of the
    per-particle search, on a specimen with weak bonds and large random
    velocities, so contacts form and bonds break.  No row may overflow."""
    def build():
        nc = native_cases.uniaxial_cylinder(ctx, 8.0, 12.0, seed=12, strength=(2e6, 4e6),
                                            skin=skin)
        N = nc.sys.N
        g = np.random.default_rng(4)
        nc.sys.set_velocities((g.standard_normal(3 * N) * 5e-2).astype(np.float32),
                              (g.standard_normal(3 * N) * 5e-2).astype(np.float32))
        return nc.sys

    a, b = build(), build()
    a.set_tiles(64, 128, True)
    b.set_tiles(64, 0, True)
    for k in range(4):
        a.step(30)
        b.step(30)
        _same_state(a, b, label=f"skin {skin} chunk {k}")
    assert a.bond_counts() == b.bond_counts()
    assert a.bond_counts()[1] > 0
    assert ctx.flags() == 0
    a.close()
    b.close()
