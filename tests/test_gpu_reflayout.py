"""GPU parity of the reference-layout tier against the CPU oracle.

Every stage is compared bit for bit (integer lists, FP32 arrays) except the
platen / FIRE reductions, which are FP64 sums of the same FP32 terms taken in
a different order (relative tolerance 1e-12).
"""
import numpy as np
import pytest

from oracle.oracle import OracleLayout, OracleUniaxialDriver
from rfd import cases
from rfd.reflayout import FLAG_CELL_RANGE, RefLayout, UniaxialDriver

pytestmark = pytest.mark.gpu

BIT_KEYS = ("R", "V", "W", "F", "M", "CI", "CIs", "pnC", "IL", "ILtype", "iL", "rij", "Oijt",
            "Mijn", "Mijt")


def assert_same(gpu: dict, cpu: dict, keys=BIT_KEYS, where=""):
    for k in keys:
        a, b = gpu[k], cpu[k]
        assert a.shape == b.shape, (where, k, a.shape, b.shape)
        if a.dtype.kind == "f":
            same = a.view(np.uint32) == b.view(np.uint32)
            if not same.all():
                bad = np.nonzero(~same)[0][:5]
                raise AssertionError(f"{where}: {k} differs at {bad}: gpu {a[bad]} cpu {b[bad]}")
        else:
            if not np.array_equal(a, b):
                bad = np.nonzero(a != b)[0][:5]
                raise AssertionError(f"{where}: {k} differs at {bad}: gpu {a[bad]} cpu {b[bad]}")


def perturb(n, seed, scale):
    g = np.random.default_rng(seed)
    return (g.standard_normal(3 * n) * scale).astype(np.float32)


def twin(ctx, case):
    grid = case.grid()
    return RefLayout(ctx, case.R, case.K, grid), OracleLayout(case.R, case.K, grid)


@pytest.mark.parametrize("D,H,seed", [(10.0, 20.0, 1), (6.0, 8.0, 7)])
def test_lists_bitexact(ctx, D, H, seed):
    case = cases.reference_uniaxial(D, H, seed=seed)
    g, o = twin(ctx, case)
    aacut = float(np.float32(case.mat.a_farcut) ** 2)
    for s in (g, o):
        s.cell_distribution()
        s.il_construct(aacut)
        s.bonds_construct(case.mat.b_dd)
        s.delete_far_links(case.mat.aa_farcut)
    assert o.last_status == 0 and ctx.flags() == 0
    assert_same(g.host(), o.host(), where="setup")
    assert (o.ILtype == 1).sum() > 4 * case.N


def test_reconstruct_after_motion_bitexact(ctx):
    case = cases.reference_uniaxial(8.0, 12.0, seed=3)
    g, o = twin(ctx, case)
    aacut = float(np.float32(case.mat.a_farcut) ** 2)
    for s in (g, o):
        s.cell_distribution()
        s.il_construct(aacut)
        s.bonds_construct(case.mat.b_dd)
    # move particles so that new pairs come inside the cut and some leave
    d = perturb(case.N, 11, 0.08)
    g.R += __import__("torch").from_numpy(d).cuda()
    o.R += d
    for s in (g, o):
        s.delete_far_links(case.mat.aa_farcut)
        s.cell_distribution()
        s.il_reconstruct_bpm(aacut, quirk_p0=True)
    assert_same(g.host(), o.host(), where="reconstruct")


def _relist_twin(ctx, R, K, grid, aacut, seed):
    """Lists built, cut back to pairs closer than 0.5, particles moved, and
    rebuilt by the reconstruction, on the GPU and in the oracle."""
    import torch

    g, o = RefLayout(ctx, R, K, grid), OracleLayout(R, K, grid)
    ctx.flags()
    for s in (g, o):
        s.cell_distribution()
        s.il_construct(aacut)
        s.delete_far_links(0.25)
    d = perturb(len(R) // 3, seed, 0.02)
    g.R += torch.from_numpy(d).cuda()
    o.R += d
    for s in (g, o):
        s.cell_distribution()
        s.il_reconstruct_bpm(aacut, quirk_p0=True)
    assert o.last_status == 0 and ctx.flags() == 0
    keys = ("R", "CI", "CIs", "pnC", "IL", "ILtype")
    assert_same(g.host(), o.host(), keys=keys, where="relist")
    return o.host()["IL"].reshape(-1, K)


def test_reconstruct_follows_caller_list_order(ctx):
    """The reconstruction inserts partners in the order of the caller's cell
    lists, whatever that order is: with every cell's members listed in
    reverse id order (CIs), the GPU's fine-grid path gives the rows of the
    oracle's 8-cell scan over the same lists, and those differ from the rows
    the id-ordered lists give."""
    import torch

    case = cases.reference_uniaxial(8.0, 12.0, seed=5)
    aacut = float(np.float32(case.mat.a_farcut) ** 2)
    rows = {}
    for order in ("id", "reversed"):
        g, o = twin(ctx, case)
        for s in (g, o):
            s.cell_distribution()
            s.il_construct(aacut)
            s.delete_far_links(0.25)
        d = perturb(case.N, 13, 0.02)
        g.R += torch.from_numpy(d).cuda()
        o.R += d
        for s in (g, o):
            s.cell_distribution()
        if order == "reversed":
            N, C = case.N, case.grid().C
            CIs = o.CIs.copy()
            for c in np.nonzero(o.pnC[C:])[0]:
                s0, n = int(o.pnC[c]), int(o.pnC[C + c])
                CIs[N + s0:N + s0 + n] = CIs[N + s0:N + s0 + n][::-1]
            o.CIs[:] = CIs
            g.CIs.copy_(torch.from_numpy(CIs.view(np.int32)).cuda())
        for s in (g, o):
            s.il_reconstruct_bpm(aacut, quirk_p0=True)
        assert o.last_status == 0 and ctx.flags() == 0
        assert_same(g.host(), o.host(), keys=("CIs", "IL", "ILtype"), where=order)
        rows[order] = o.host()["IL"]
    changed = (rows["id"] != rows["reversed"]).sum()
    print(f"slots placed differently by the reversed lists: {changed}", flush=True)
    assert changed > 0


def test_reconstruct_dense_cluster_bitexact(ctx):
    """50 particles packed in a ball of radius 0.6 inside a lattice, so that
    each has 49 partners in reach: more than the 32 the fine-grid search
    sorts, which sends them through the 8-cell scan."""
    from rfd.reflayout import Grid

    rng = np.random.default_rng(21)
    ax = np.arange(1.5, 15.0, 1.0)
    lat = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    c0 = np.array([7.3, 7.3, 7.3])
    lat = lat[np.linalg.norm(lat - c0, axis=1) > 2.5]
    u = rng.standard_normal((50, 3))
    ball = c0 + u / np.linalg.norm(u, axis=1, keepdims=True) * 0.6 * rng.uniform(0, 1, (50, 1)) ** (1 / 3)
    pos = np.concatenate([lat, ball])
    R = np.ascontiguousarray(pos.T.reshape(-1)).astype(np.float32)
    a = np.float32(2.9)
    grid = Grid(float(a), float(np.float32(1.0) / a), 6, 6, 6)
    aacut = float(np.float32(1.45) ** 2)
    IL = _relist_twin(ctx, R, 64, grid, aacut, seed=22)
    full = (IL != 0xFFFFFFFF).sum(axis=1)
    assert full[len(lat):].min() > 32


def test_reconstruct_wide_grid_bitexact(ctx):
    """A grid of 25,000 x 1 x 1 cells, too long for the fine binning (more
    than 40,000 fine cells along x): the reconstruction scans the 8 cells."""
    from rfd.reflayout import Grid

    rng = np.random.default_rng(31)
    n = 300
    pos = np.stack([rng.uniform(50.0, 80.0, n), rng.uniform(0.2, 2.7, n),
                    rng.uniform(0.2, 2.7, n)], -1)
    R = np.ascontiguousarray(pos.T.reshape(-1)).astype(np.float32)
    a = np.float32(2.9)
    grid = Grid(float(a), float(np.float32(1.0) / a), 25000, 1, 1)
    aacut = float(np.float32(1.45) ** 2)
    IL = _relist_twin(ctx, R, 32, grid, aacut, seed=32)
    assert (IL != 0xFFFFFFFF).sum() > 5 * n


def test_forces_bitexact_with_motion(ctx):
    import torch

    case = cases.reference_uniaxial(8.0, 12.0, seed=5)
    g, o = twin(ctx, case)
    m = case.mat
    aacut = float(np.float32(m.a_farcut) ** 2)
    for s in (g, o):
        s.cell_distribution()
        s.il_construct(aacut)
        s.bonds_construct(m.b_dd)
        s.delete_far_links(m.aa_farcut)
    dv = perturb(case.N, 21, 1e-3)
    dw = perturb(case.N, 22, 2e-3)
    for k in range(12):
        g.V.copy_(torch.from_numpy(dv * (k + 1)).cuda())
        g.W.copy_(torch.from_numpy(dw * (k + 1)).cuda())
        o.V[:] = dv * (k + 1)
        o.W[:] = dw * (k + 1)
        for s in (g, o):
            s.zero_fm()
            s.forces_bpm(m)
            s.forces_contact(m)
            s.increments(m)
        assert_same(g.host(), o.host(), where=f"iteration {k}")
    assert np.abs(o.F).max() > 0


@pytest.mark.parametrize("K", [32, 27])
def test_contact_forces_dense_bitexact(ctx, K):
    """Every neighbour pair a touching contact: the lattice cylinder shrunk by
    3 % about its centre, lists built without bonds, so each particle touches
    about 12 others. At K = 32 a workgroup of the four-slot contact kernel
    then holds over 256 touching slots and runs the law and the row sums in
    rounds; K = 27 takes the one-slot kernel."""
    import torch

    case = cases.reference_uniaxial(8.0, 12.0, seed=6, K=K)
    N, m = case.N, case.mat
    R = case.R.reshape(3, N).astype(np.float64)
    c = R.mean(axis=1, keepdims=True)
    R = np.ascontiguousarray((c + (R - c) * 0.97).reshape(-1)).astype(np.float32)
    grid = case.grid()
    g, o = RefLayout(ctx, R, K, grid), OracleLayout(R, K, grid)
    for s in (g, o):
        s.cell_distribution()
        s.il_construct(float(np.float32(m.a_farcut) ** 2))
    dv, dw = perturb(N, 61, 1e-3), perturb(N, 62, 2e-3)
    for it in range(3):
        g.V.copy_(torch.from_numpy(dv * (it + 1)).cuda())
        g.W.copy_(torch.from_numpy(dw * (it + 1)).cuda())
        o.V[:] = dv * (it + 1)
        o.W[:] = dw * (it + 1)
        for s in (g, o):
            s.zero_fm()
            s.forces_contact(m)
        assert_same(g.host(), o.host(), where=f"K={K} iteration {it}")
    # touching slots per 1024-slot group (32 rows at K = 32)
    IL = o.IL.reshape(N, K)
    Rr = R.reshape(3, N)
    live = np.zeros(N, np.int64)
    for k in range(K):
        j = IL[:, k]
        ok = j != 0xFFFFFFFF
        d = np.linalg.norm(Rr[:, j[ok]] - Rr[:, ok], axis=0)
        live[ok] += d < 2 * m.p_r
    per_group = np.add.reduceat(live, np.arange(0, N, 1024 // K))
    print(f"K={K}: touching per particle {live.mean():.1f}, per group max {per_group.max()}")
    assert per_group.max() > 256 and ctx.flags() == 0


@pytest.mark.parametrize("k", [1, 2, 5, 64])
def test_forces_short_and_long_rows_bitexact(ctx, k):
    """The force kernels at row lengths far from 32: the rows of a K = 32
    bond construction cut to their first k slots (k = 1 puts 256 rows in a
    bond workgroup, one per thread) or padded with empty slots to 64, then
    a few force passes with growing motion that breaks most bonds."""
    import torch

    case = cases.reference_uniaxial(10.0, 14.0, seed=12)  # 1221 particles: 5 workgroups at k = 1
    m = case.mat
    base = OracleLayout(case.R, case.K, case.grid())
    base.cell_distribution()
    base.il_construct(float(np.float32(m.a_farcut) ** 2))
    base.bonds_construct(m.b_dd)
    N, K0 = case.N, case.K

    def rows(a, width):
        a = a.reshape(N, K0, width)
        if k <= K0:
            return a[:, :k].reshape(-1)
        pad = np.zeros((N, k - K0, width), a.dtype)
        if a.dtype == np.uint32:
            pad[:] = 0xFFFFFFFF
        return np.concatenate([a, pad], axis=1).reshape(-1)

    g, o = RefLayout(ctx, case.R, k, case.grid()), OracleLayout(case.R, k, case.grid())
    o.IL[:] = rows(base.IL, 1)
    o.ILtype[:] = rows(base.ILtype, 1)
    o.iL[:] = rows(base.iL, 1)
    for name in ("rij", "Oijt", "Mijn", "Mijt"):
        o.slot[name][:] = rows(base.slot[name], 3)
        g.slot[name].copy_(torch.from_numpy(o.slot[name]).cuda())
    g.IL.copy_(torch.from_numpy(o.IL.view(np.int32)).cuda())
    g.ILtype.copy_(torch.from_numpy(o.ILtype).cuda())
    g.iL.copy_(torch.from_numpy(o.iL).cuda())
    n0 = int((o.ILtype == 1).sum())
    assert n0 > 0
    dv = perturb(N, 41, 2e-2)
    dw = perturb(N, 42, 2e-2)
    for it in range(4):
        g.V.copy_(torch.from_numpy(dv * (it + 1)).cuda())
        g.W.copy_(torch.from_numpy(dw * (it + 1)).cuda())
        o.V[:] = dv * (it + 1)
        o.W[:] = dw * (it + 1)
        for s in (g, o):
            s.zero_fm()
            s.forces_bpm(m)
            s.forces_contact(m)
            s.increments(m)
        assert_same(g.host(), o.host(), where=f"k={k} iteration {it}")
    n1 = int((o.ILtype == 1).sum())
    print(f"k={k}: bond slots {n0} -> {n1}", flush=True)
    assert 0 < n1 < n0


def test_bond_break_bitexact(ctx):
    """Weak bonds break under the same motion on both sides."""
    import dataclasses

    import torch

    case = cases.reference_uniaxial(8.0, 12.0, seed=9)
    m = dataclasses.replace(case.mat, sc=1e-4, tc=1e-4)
    g, o = twin(ctx, case)
    aacut = float(np.float32(m.a_farcut) ** 2)
    for s in (g, o):
        s.cell_distribution()
        s.il_construct(aacut)
        s.bonds_construct(m.b_dd)
    dv = perturb(case.N, 31, 2e-2)
    g.V.copy_(torch.from_numpy(dv).cuda())
    o.V[:] = dv
    before = int((o.ILtype == 1).sum())
    for k in range(5):
        for s in (g, o):
            s.zero_fm()
            s.forces_bpm(m)
            s.forces_contact(m)
            s.increments(m)
    assert_same(g.host(), o.host(), where="break")
    assert int((o.ILtype == 1).sum()) < before


@pytest.mark.parametrize("mode", ["uniaxial", "brazil"])
def test_driver_trajectory_bitexact(ctx, mode):
    """Whole explicit steps (cell sort, reconstruct, forces, platens,
    integration, cut, prune) for 60 steps; state compared every 20."""
    mk = cases.reference_uniaxial if mode == "uniaxial" else cases.reference_brazil
    args = (8.0, 14.0) if mode == "uniaxial" else (14.0, 5.0)
    cg, co = mk(*args, seed=2), mk(*args, seed=2)
    # faster loading so that the platens engage within the test
    for c in (cg, co):
        c.cmp.v_load *= 50
    dg = UniaxialDriver(ctx, cg.R, cg.mat, cg.box, cg.cmp, cg.K)
    do = OracleUniaxialDriver(co.R, co.mat, co.box, co.cmp, co.grid(), co.K)
    dg.setup()
    do.setup()
    for k in range(60):
        sg = dg.step()
        so = do.step()
        np.testing.assert_allclose(sg, so, rtol=1e-12, atol=1e-300)
        if k % 20 == 19:
            assert_same(dg.s.host(), do.s.host(), where=f"{mode} step {k}")
    assert ctx.flags() == 0
    assert cg.cmp.stress == pytest.approx(co.cmp.stress, rel=1e-12)


@pytest.mark.parametrize("K", [24, 27, 40])
def test_driver_trajectory_other_row_lengths(ctx, K):
    """The same whole steps with K slots per particle other than 32: the
    contact kernel then tiles floor(256 / K) particles per workgroup with idle
    threads at the end (K = 27: 9 particles, 243 of 256 threads), the bond
    kernel floor(512 / K) (18 rows, 486 slots), and the list
    reconstruction pads its LDS rows to 28, 28 and 44 words."""
    args = (8.0, 14.0)
    cg = cases.reference_uniaxial(*args, seed=5, K=K)
    co = cases.reference_uniaxial(*args, seed=5, K=K)
    for c in (cg, co):
        c.cmp.v_load *= 50
    dg = UniaxialDriver(ctx, cg.R, cg.mat, cg.box, cg.cmp, cg.K)
    do = OracleUniaxialDriver(co.R, co.mat, co.box, co.cmp, co.grid(), co.K)
    dg.setup()
    do.setup()
    for k in range(40):
        np.testing.assert_allclose(dg.step(), do.step(), rtol=1e-12, atol=1e-300)
        if k % 20 == 19:
            assert_same(dg.s.host(), do.s.host(), where=f"K={K} step {k}")
    assert ctx.flags() == 0 and do.s.last_status == 0
    assert (do.s.ILtype == 1).sum() > 4 * cg.N


def test_fire_kernels(ctx):
    import torch

    case = cases.reference_uniaxial(6.0, 8.0, seed=4)
    g, o = twin(ctx, case)
    f = perturb(case.N, 41, 1e-2)
    v = perturb(case.N, 42, 1e-3)
    g.F.copy_(torch.from_numpy(f).cuda())
    g.V.copy_(torch.from_numpy(v).cuda())
    o.F[:] = f
    o.V[:] = v
    pg, po = g.fire_power(), o.fire_power()
    np.testing.assert_allclose(pg, po, rtol=1e-12)
    lib = ctx.lib
    from rfd.reflayout import _p
    from oracle.oracle import lib as olib, ptr

    ctx.call("rfd_fire_increments", _p(g.F), _p(g.V), _p(g.R), case.N, 0.01, 0.02, 0.1)
    olib().orc_fire_increments(ptr(o.F), ptr(o.V), ptr(o.R), case.N, 0.01, 0.02, 0.1)
    ctx.call("rfd_fire_halfstep", _p(g.V), _p(g.R), case.N, 0.005)
    olib().orc_fire_halfstep(ptr(o.V), ptr(o.R), case.N, 0.005)
    assert_same(g.host(), o.host(), keys=("R", "V"), where="fire")
    del lib


def test_cell_distribution_many_chunks_bitexact(ctx):
    """Cell binning on a grid of 43,200 cells (the scan runs in chunks of 8,192
    cells, so this spans six of them, the last one partial): fewer particles
    than cells, so most cells are empty and some hold several."""
    from rfd.reflayout import Grid

    rng = np.random.default_rng(11)
    N = 20000
    hi = np.array([20.0, 20.0, 13.5])
    pos = rng.uniform(0.0, 1.0, (N, 3)) * hi * 0.999
    R = np.ascontiguousarray(pos.T.reshape(-1)).astype(np.float32)
    grid = Grid(0.5, float(np.float32(1.0 / np.float32(0.5))), 40, 40, 27)
    assert grid.C > 5 * 8192
    g, o = RefLayout(ctx, R, 4, grid), OracleLayout(R, 4, grid)
    ctx.flags()
    g.cell_distribution()
    o.cell_distribution()
    assert o.last_status == 0 and ctx.flags() == 0
    assert_same(g.host(), o.host(), keys=("CI", "CIs", "pnC"), where="cells")
    counts = o.host()["pnC"][grid.C:]
    assert counts.sum() == N and (counts == 0).sum() > grid.C // 2 and counts.max() > 1


@pytest.mark.parametrize("crowd", [60, 65, 200])
def test_cell_distribution_crowded_cells_bitexact(ctx, crowd):
    """Cells holding up to 64 particles are ranked by id within one wave's
    lanes, larger ones by the general loop: `crowd` particles in one cell and
    crowd - 1 in another (plus the sparse ones that fall there: 63 and 61,
    67 and 65, 203 and 202 members), in shuffled id order."""
    from rfd.reflayout import Grid

    rng = np.random.default_rng(crowd)
    sparse = rng.uniform(0.0, 1.0, (3000, 3)) * np.array([10.0, 10.0, 10.0]) * 0.999
    a = rng.uniform(2.05, 2.45, (crowd, 3))           # inside cell (2, 2, 2) of side 1
    b = rng.uniform(7.05, 7.45, (crowd - 1, 3))       # inside cell (7, 7, 7)
    pos = np.concatenate([sparse, a, b])[rng.permutation(3000 + 2 * crowd - 1)]
    R = np.ascontiguousarray(pos.T.reshape(-1)).astype(np.float32)
    grid = Grid(1.0, 1.0, 10, 10, 10)
    g, o = RefLayout(ctx, R, 4, grid), OracleLayout(R, 4, grid)
    ctx.flags()
    g.cell_distribution()
    o.cell_distribution()
    assert o.last_status == 0 and ctx.flags() == 0
    assert_same(g.host(), o.host(), keys=("CI", "CIs", "pnC"), where=f"crowd {crowd}")
    assert o.host()["pnC"][grid.C:].max() >= crowd


def test_cell_range_flag(ctx):
    """A particle outside the grid raises RFD_FLAG_CELL_RANGE, not a fault."""
    import torch

    case = cases.reference_uniaxial(6.0, 8.0, seed=4)
    g, _ = twin(ctx, case)
    ctx.flags()
    g.R[0] = -5.0
    g.cell_distribution()
    torch.cuda.synchronize()
    assert ctx.flags() & FLAG_CELL_RANGE


def test_standalone_sumup_and_break_bitexact(ctx):
    """The stand-alone sum-up (md_sumupforcesdem.cu:19) and break check
    (md_checkbreakconditiondem.cu:20) on the slot arrays of a bonded specimen
    with random slot loads: F, M, ILtype and Oijt bit for bit."""
    import torch

    case = cases.reference_uniaxial(6.0, 8.0, seed=5)
    g, o = twin(ctx, case)
    m = case.mat
    for s in (g, o):
        s.cell_distribution()
        s.il_construct(float(np.float32(m.a_farcut) ** 2))
        s.bonds_construct(m.b_dd)
    assert_same(g.host(), o.host(), where="setup")
    rng = np.random.default_rng(17)
    S = case.N * case.K
    scale = {"rij": 1.0, "Oijt": 1e-3, "Fijn": 2e-3, "Fijt": 1e-3, "Mijn": 1e-4,
             "Mijt": 1e-4, "Mijadd": 1e-5}
    for k, sc in scale.items():
        a = (rng.standard_normal(3 * S) * sc).astype(np.float32)
        o.slot[k][:] = a
        g.slot[k].copy_(torch.from_numpy(a).cuda())
    for s in (g, o):
        s.sumup_forces()
    # thresholds near the median stress so that a good share of bonds break
    A = float(np.float32(3.14159265358979)) * m.b_r ** 2
    sc_, tc_ = 2e-3 / A, 1.5e-3 / A
    for s in (g, o):
        s.check_break(sc_, tc_, m.b_r)
    hg, ho = g.host(), o.host()
    assert_same(hg, ho, keys=("F", "M", "ILtype", "Oijt"), where="sumup/break")
    alive = int((ho["ILtype"] == 1).sum())
    assert 0 < alive < int((o.IL != 0xFFFFFFFF).sum())  # some, not all, slots broke


def test_brazil_platen_and_cut_bitexact(ctx):
    """Brazilian platens (md_braziltest_simple2.cu:19) with in-plane friction,
    then the cut box (md_parallelepipedrestriction.cu:73), on a disc with
    random velocities and forces: R, V, F bit for bit, reaction sums to 1e-12."""
    import torch

    case = cases.reference_brazil(10.0, 3.0, seed=6)
    g, o = twin(ctx, case)
    rng = np.random.default_rng(23)
    f = (rng.standard_normal(3 * case.N) * 1e-2).astype(np.float32)
    v = (rng.standard_normal(3 * case.N) * 1e-3).astype(np.float32)
    for s in (g, o):
        if s is g:
            g.F.copy_(torch.from_numpy(f).cuda())
            g.V.copy_(torch.from_numpy(v).cuda())
        else:
            o.F[:] = f
            o.V[:] = v
    cmp = case.cmp
    zcut = 1e3
    sg, so = g.platen_brazil(cmp, zcut), o.platen_brazil(cmp, zcut)
    np.testing.assert_allclose(sg, so, rtol=1e-12)
    assert sg[0] < 0 < sg[1]
    c = case.box.center
    half = [3.0, 3.0, 3.0]  # cuts the outer part of the 10 mm disc
    hp = case.box.hidden_point
    for s in (g, o):
        s.cut_restriction(c, half, hp)
    hg, ho = g.host(), o.host()
    assert_same(hg, ho, keys=("R", "V", "F"), where="brazil/cut")
    assert np.any(ho["R"] == np.float32(hp[0]))
