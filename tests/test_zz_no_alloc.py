"""rfd_cell_distribution_noalloc against rfd_cell_distribution through the C
ABI: the same cell keys, sorted lists and cell table bit for bit, on two calls
in a row (the second with fewer particles and a smaller grid, so the fixed
workspace still holds the first call's larger contents), and both against the
oracle.  The capacity check rejects an input beyond the fixed workspace."""
import numpy as np
import pytest
import torch

from oracle.oracle import OracleLayout
from rfd._lib import RfdError
from rfd.reflayout import Grid, RefLayout, _p

pytestmark = pytest.mark.gpu

RFD_E_CAPACITY = -4


def _case(seed, N, dims, a):
    rng = np.random.default_rng(seed)
    hi = np.array(dims, np.float64) * a * 0.999
    R = np.ascontiguousarray((rng.uniform(0.0, 1.0, (N, 3)) * hi).T.reshape(-1)).astype(np.float32)
    grid = Grid(a, float(np.float32(1.0 / np.float32(a))), *dims)
    return R, grid


def _noalloc(ctx, g, out):
    gr = g.grid
    ctx.call("rfd_cell_distribution_noalloc", _p(g.R), g.N, gr.inv_a, gr.Nx, gr.Ny, gr.Nz,
             _p(out["CI"]), _p(out["CIs"]), _p(out["pnC"]))


@pytest.mark.parametrize("first,second", [((30000, (40, 40, 27)), (5000, (12, 10, 9)))])
def test_noalloc_matches_allocating_entry(ctx, first, second):
    for seed, (N, dims) in enumerate((first, second)):
        print(f"[noalloc] call {seed + 1}: N={N} grid={dims}", flush=True)
        R, grid = _case(seed + 21, N, dims, 0.5)
        g = RefLayout(ctx, R, 4, grid)
        ctx.flags()
        g.cell_distribution()
        ref = {k: getattr(g, k).clone() for k in ("CI", "CIs", "pnC")}
        # outputs start as garbage: every word must be written by the call
        out = {k: torch.full_like(v, 0x5A5A5A5A) for k, v in ref.items()}
        _noalloc(ctx, g, out)
        ctx.sync()
        assert ctx.flags() == 0
        for k in ref:
            assert torch.equal(out[k], ref[k]), (seed, k)
        o = OracleLayout(R, 4, grid)
        o.cell_distribution()
        assert o.last_status == 0
        for k in ("CI", "CIs", "pnC"):
            assert np.array_equal(out[k].cpu().numpy().view(np.uint32), getattr(o, k)), (seed, k)


def test_noalloc_rejects_input_beyond_fixed_workspace(ctx):
    """16,777,216 words hold N + C + ceil(C / 8192); one particle more than
    fits is refused before anything runs."""
    R, grid = _case(3, 8, (4, 4, 4), 0.5)
    g = RefLayout(ctx, R, 4, grid)
    C = grid.C
    N_over = (16 << 20) - C  # with the one chunk-sum word: one word too many
    with pytest.raises(RfdError) as e:
        ctx.call("rfd_cell_distribution_noalloc", _p(g.R), N_over, grid.inv_a, grid.Nx,
                 grid.Ny, grid.Nz, _p(g.CI), _p(g.CIs), _p(g.pnC))
    assert e.value.status == RFD_E_CAPACITY
