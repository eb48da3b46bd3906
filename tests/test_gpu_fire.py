"""Quasi-static load steps (axial strain increment + FIRE) on the native tier.

* every load step converges to |F|max < tol F_ref (F_ref = E pi r^2);
* the elastic response is linear: after a seating preload, equal strain
  increments add equal reactions (2 %; the reaction error of a relaxed step
  is bounded by the interior residual, N tol F_ref, since the slowest FIRE
  mode is the coherent axial one, so this test relaxes to tol = 1e-8);
* a warm start from the previous step's increment reaches the same residual
  in fewer FIRE iterations, and the relaxed reactions agree (0.5 %);
* a step capped at one iteration reports non-convergence (the fallback path
  of the learned warm start keys on this).
"""
import numpy as np
import pytest

from rfd import native, native_cases

pytestmark = pytest.mark.gpu


def _system(ctx, seed=5, rel_tol=1e-6):
    nc = native_cases.uniaxial_cylinder(ctx, 8.0, 12.0, seed=seed, strength=(1e12, 1e12))
    nc.sys.seat_platens(preload=0.05)
    fp = native.fire_params(nc.mat.E, 0.5, nc.mat.dt, rel_tol=rel_tol, max_iter=50000,
                            check_every=200)
    return nc, fp


def test_fire_load_steps_converge_linearly(ctx):
    nc, fp = _system(ctx, rel_tol=1e-8)
    s = nc.sys
    res = [s.fire_load_step(-1e-4, fp) for _ in range(5)]
    bound = s.N * fp.ftol
    for r in res:
        assert r.converged, r.as_dict()
        assert r.fmax < fp.ftol
    top = np.array([r.sum_top for r in res])
    inc = np.diff(top)
    assert top[0] > 0 and inc[0] > 50 * bound
    np.testing.assert_allclose(inc / inc[0], np.ones_like(inc), rtol=0.02)
    fr, mr = s.residual()
    assert fr < fp.ftol * 1.01
    assert ctx.flags() == 0


def test_warm_start_saves_iterations_at_same_residual(ctx):
    (na, fp), (nb, _) = _system(ctx, seed=7), _system(ctx, seed=7)
    for _ in range(3):
        ra = na.sys.fire_load_step(-1e-5, fp)
        rb = nb.sys.fire_load_step(-1e-5, fp)
    du, _, dth, _ = nb.sys.fire_buffers()
    guess_u, guess_t = du.clone(), dth.clone()
    ra = na.sys.fire_load_step(-1e-5, fp)
    rb = nb.sys.fire_load_step(-1e-5, fp, guess_u, guess_t)
    assert ra.converged and rb.converged
    assert rb.fmax0 < ra.fmax0
    assert rb.iterations < ra.iterations, (ra.as_dict(), rb.as_dict())
    assert rb.sum_top == pytest.approx(ra.sum_top, rel=5e-3)


def test_iteration_cap_reports_non_convergence(ctx):
    nc, fp = _system(ctx, seed=9)
    fp.max_iter = 1
    fp.check_every = 1
    r = nc.sys.fire_load_step(-1e-5, fp)
    assert not r.converged and r.iterations <= 1


def test_measure_fused_into_force_pass_matches_separate_measure(ctx):
    """The FIRE measure in the force pass's epilogue (fusion on, the default)
    against the separate measure kernel (fusion off), from the same state:
    the two differ only in the order of the FP64 partial sums, so the load
    steps converge alike (iterations within 5 %) to the same reactions (1e-4
    relative, the spread of where two relaxations stop below tol) and the
    same positions (1e-4 of the step's displacement)."""
    (na, fp), (nb, _) = _system(ctx, seed=11, rel_tol=1e-8), _system(ctx, seed=11, rel_tol=1e-8)
    a, b = na.sys, nb.sys
    b.set_fusion(False)
    assert a.fused() and not b.fused()
    R0 = a.state()["R"].copy()
    for k in range(3):
        ra = a.fire_load_step(-1e-4, fp)
        rb = b.fire_load_step(-1e-4, fp)
        print(f"[fire fused measure] step {k}: iterations {ra.iterations} / {rb.iterations}, "
              f"top {ra.sum_top:.9e} / {rb.sum_top:.9e}", flush=True)
        assert ra.converged and rb.converged
        # the first measure of a step runs separately in both: the same bits
        # from the same state, then close
        if k == 0:
            assert ra.fmax0 == rb.fmax0
        assert ra.fmax0 == pytest.approx(rb.fmax0, rel=1e-4)
        assert abs(ra.iterations - rb.iterations) <= 0.05 * rb.iterations + 5
        assert ra.sum_top == pytest.approx(rb.sum_top, rel=1e-4)
        assert ra.sum_bottom == pytest.approx(rb.sum_bottom, rel=1e-4)
    Ra, Rb = a.state()["R"], b.state()["R"]
    assert np.abs(Ra - Rb).max() <= 1e-4 * np.abs(Rb - R0).max()
    assert ctx.flags() == 0
