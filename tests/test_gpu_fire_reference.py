"""The reference FIRE loop as a library driver (rfd_fire_relax_reference,
md_firerelaxation.cu:201-304) against the oracle's restatement of the same
loop (oracle.fire_relax_reference), both driving their own copies of the same
force field: positions, velocities and the loop state (float dt and alpha,
counters) must agree bit for bit.  The power is summed in FP64 on both sides
(in a different order), so the two could only part where it is zero to
rounding; the state at every call shows they do not."""
import numpy as np
import pytest
import torch

from oracle.oracle import FireRefState, OracleLayout, fire_relax_reference
from rfd import cases
from rfd.reflayout import RefLayout, ReferenceFire

pytestmark = pytest.mark.gpu


def _same_state(d: ReferenceFire, o: FireRefState):
    s = d.state
    assert (s.steps, s.n_positive, s.n_negative, bool(s.stopped)) == \
        (o.steps, o.n_positive, o.n_negative, o.stopped)
    assert np.float32(s.dt) == o.dt and np.float32(s.alpha) == o.alpha


def test_reference_fire_quadratic_bowl_bitexact(ctx):
    """F = -k (R - R0) computed on each side in FP32 (same operations)."""
    from rfd.reflayout import Grid

    g = np.random.default_rng(7)
    N = 4096
    R0 = g.uniform(0, 5, 3 * N).astype(np.float32)
    R = (R0 + g.normal(0, 0.1, 3 * N)).astype(np.float32)
    k = np.float32(2.0)
    lay = RefLayout(ctx, R, 1, Grid(1, 1, 1, 1, 1))
    R0d = torch.from_numpy(R0).cuda()
    kd = torch.tensor(k, device="cuda")

    def dev_forces():
        torch.mul(R0d - lay.R, kd, out=lay.F)

    o = OracleLayout(R, 1, Grid(1, 1, 1, 1, 1))

    def orc_forces():
        o.F[:] = (R0 - o.R) * k

    d, f = ReferenceFire(0.05, 0.25), FireRefState(0.05, 0.25)
    for chunk in (1, 20, 200, 500):
        d.run(lay, 0.05, chunk, dev_forces)
        fire_relax_reference(o.F, o.V, o.R, N, 0.05, f, chunk, orc_forces)
        _same_state(d, f)
        h = lay.host()
        assert np.array_equal(h["R"].view(np.uint32), o.R.view(np.uint32))
        assert np.array_equal(h["V"].view(np.uint32), o.V.view(np.uint32))
    assert np.abs(o.R - R0).max() < 1e-4


def test_reference_fire_bonded_specimen_bitexact(ctx):
    """The bonded reference cylinder D6 x H8 (bond + contact forces of the
    slot layout, d_CalculateForcesDEM_22/_21) shaken and relaxed by the
    reference loop for 600 iterations, with the cylinder restriction role
    played by the cut box: the dt growth drives the stiff bond modes past
    stability, so the loop backs off and restarts many times."""
    case = cases.reference_uniaxial(6.0, 8.0, seed=3)
    m = case.mat
    grid = case.grid()
    g_, o = RefLayout(ctx, case.R, case.K, grid), OracleLayout(case.R, case.K, grid)
    aacut = float(np.float32(m.a_farcut) ** 2)
    for s in (g_, o):
        s.cell_distribution()
        s.il_construct(aacut)
        s.bonds_construct(m.b_dd)
    jit = (np.random.default_rng(5).standard_normal(3 * case.N) * 2e-3).astype(np.float32)
    o.R += jit
    g_.R.copy_(torch.from_numpy(o.R.copy()).cuda())
    box = case.box

    def forces(s):
        def fn():
            s.zero_fm()
            s.forces_bpm(m)
            s.forces_contact(m)
        return fn

    def restrict(s):
        def fn():
            s.cut_restriction(box.center, box.spacesized2, box.hidden_point)
        return fn

    d, f = ReferenceFire(m.dt, 0.25), FireRefState(m.dt, 0.25)
    backoffs = 0
    for _ in range(60):
        d.run(g_, m.dt_d_m, 10, forces(g_), restrict(g_))
        fire_relax_reference(o.F, o.V, o.R, case.N, m.dt_d_m, f, 10, forces(o), restrict(o))
        _same_state(d, f)
        backoffs += f.n_negative > 0
    h = g_.host()
    for k in ("R", "V", "F", "Oijt", "Mijn", "Mijt", "ILtype"):
        a, b = h[k], (o.slot[k] if k in o.slot else getattr(o, k))
        if a.dtype.kind == "f":
            a, b = a.view(np.uint32), b.view(np.uint32)
        assert np.array_equal(a, b), k
    print(f"[fire reference] steps {f.steps}, dt {f.dt:.3e} (dt0 {m.dt:.3e}), "
          f"calls ending on an uphill run {backoffs}", flush=True)
    assert backoffs > 0 and ctx.flags() == 0
