"""The quasi-static FIRE load step (rfd_system_fire_load_step) against
mechanics and against the CPU FP64 restatement of the step
(oracle/native_fire.py).

* a straight bonded chain between the uniaxial platens, compressed by one
  strain step: at equilibrium every bond carries E A eps with eps the applied
  strain (bonds in series), the top platen reads that force, and the
  particles sit on the line between the platens at equal spacing;
* a small lattice cylinder on the FP64 path, three load steps: iteration
  counts equal, residuals, reactions and positions within 1e-9 relative, the
  same bonds intact; then with weak bonds and a large step, so that bonds
  break at equilibrium and the step relaxes again (relax-and-break rounds).
"""
import numpy as np
import pytest

from oracle.native_fire import OracleFireLoadStep
from rfd import native, native_cases
from test_gpu_native_oracle import oracle_of

pytestmark = pytest.mark.gpu


def _chain(ctx, n, fp64):
    """n spheres of radius 0.5 on the z axis at spacing 1, bonds of radius
    0.4 between neighbours, platens on the end spheres."""
    mat = native_cases.material_c(0.4)
    R = np.zeros((n, 3))
    R[:, 0] = R[:, 1] = 5.0
    R[:, 2] = 5.0 + np.arange(n, dtype=np.float64)
    p = native.params_from_material(mat, refmode=0, fp64=fp64)
    p.farcut = 1.45
    s = native.NativeSystem(ctx, p, R, np.full(n, 0.5, np.float32))
    bi = np.arange(n - 1, dtype=np.uint32)
    s.set_bonds(native.Bonds(bi, bi + 1, np.full(n - 1, 0.4, np.float32),
                             np.full(n - 1, 1e3, np.float32), np.full(n - 1, 1e3, np.float32)))
    L = native.Loading()
    L.mode = 1
    L.cx = L.cy = 5.0
    L.RR = 4.0
    L.zcut = 1e30
    L.bottom, L.top = 5.0, 5.0 + (n - 1)
    L.band = 0.5
    L.top_r, L.bottom_r = L.top - 0.5, L.bottom + 0.25
    L.size = L.size0 = L.top - L.bottom
    L.inv_area = 1.0
    s.set_loading(L)
    return s, mat


@pytest.mark.parametrize("fp64", [False, True])
def test_bonded_chain_relaxes_to_series_springs(ctx, fp64):
    n = 12
    s, mat = _chain(ctx, n, fp64)
    fp = native.fire_params(mat.E, 0.5, mat.dt, rel_tol=1e-10, max_iter=200000,
                            check_every=500)
    eps = -2e-4
    r = s.fire_load_step(eps, fp)
    assert r.converged and r.fmax < fp.ftol, r.as_dict()
    A = np.pi * 0.4 * 0.4
    Fbond = mat.E * A * abs(eps)          # every bond of the series carries it
    # reactions: both platens hold the chain's force (compression)
    assert r.sum_top == pytest.approx(Fbond, rel=1e-4)
    assert r.sum_bottom == pytest.approx(-Fbond, rel=1e-4)
    z = s.state()["R"].reshape(3, n).T[:, 2]
    want = 5.0 + np.arange(n) * (1.0 + eps)
    np.testing.assert_allclose(z, want, rtol=0, atol=1e-8)
    print(f"[chain fp64={fp64}] iterations {r.iterations}, reaction {r.sum_top:.6e} vs "
          f"{Fbond:.6e}", flush=True)


def _compare(r, ro, s, o, tol=1e-9):
    """Iterations and break counts equal; reactions within tol relative;
    residuals (max |F_i|, a difference of forces of order F_ref) within
    1e-12 F_ref, F_ref = E pi r_mean^2; positions within tol of their size."""
    assert r.iterations == ro["iterations"], (r.as_dict(), ro)
    assert bool(r.converged) == ro["converged"]
    assert (r.breaks, r.break_rounds) == (ro["breaks"], ro["break_rounds"])
    fref = s.params.E * np.pi * 0.25
    for k in ("fmax0", "fmax"):
        a, b = getattr(r, k), ro[k]
        assert abs(a - b) <= 1e-12 * fref, (k, a, b)
    for k in ("sum_top", "sum_bottom"):
        a, b = getattr(r, k), ro[k]
        assert a == pytest.approx(b, rel=tol, abs=1e-12 * fref), (k, a, b)
    st = s.state64()
    scale = np.abs(o.R).max()
    assert np.abs(st["R"] - o.R).max() <= tol * scale
    fl = s.bond_state()[1]
    assert np.array_equal(fl == 1, o.act == 1)


def test_fp64_load_steps_match_cpu_restatement(ctx):
    nc = native_cases.uniaxial_cylinder(ctx, 4.0, 6.0, seed=8, strength=(1e12, 1e12),
                                        fp64=True, flat_ends=True)
    s = nc.sys
    s.seat_platens(preload=0.05)
    fp = native.fire_params(nc.mat.E, 0.5, nc.mat.dt, rel_tol=1e-8, max_iter=50000,
                            check_every=100)
    o = oracle_of(s)
    of = OracleFireLoadStep(o, fp)
    for k in range(3):
        r = s.fire_load_step(-1e-4, fp)
        ro = of.load_step(-1e-4)
        print(f"[fire fp64] step {k}: N={s.N} iterations {r.iterations} / {ro['iterations']}, "
              f"reaction {r.sum_top:.6e} / {ro['sum_top']:.6e}", flush=True)
        _compare(r, ro, s, o)


def test_fp64_load_step_with_breaks_matches_cpu_restatement(ctx):
    """Bond strengths N(20, 4) / N(40, 8) MPa, platens seated 2 um into the
    end layers and a strain step of -2e-3: bonds break when judged at
    equilibrium and the step relaxes again until none breaks; three such
    steps."""
    nc = native_cases.uniaxial_cylinder(ctx, 4.0, 6.0, seed=9, strength=(20e6, 40e6),
                                        fp64=True, flat_ends=True)
    s = nc.sys
    s.seat_platens(preload=0.002)
    fp = native.fire_params(nc.mat.E, 0.5, nc.mat.dt, rel_tol=1e-8, max_iter=50000,
                            check_every=100)
    o = oracle_of(s)
    of = OracleFireLoadStep(o, fp)
    rounds = 0
    for k in range(3):
        r = s.fire_load_step(-2e-3, fp)
        ro = of.load_step(-2e-3)
        print(f"[fire fp64 breaks] step {k}: N={s.N} iterations {r.iterations} / "
              f"{ro['iterations']}, breaks {r.breaks} / {ro['breaks']} in {r.break_rounds} / "
              f"{ro['break_rounds']} rounds", flush=True)
        _compare(r, ro, s, o)
        rounds += r.break_rounds
    assert rounds >= 2
