"""Checkpoint / resume of the reference-layout tier through DATA.dat on the
GPU: a run continued from the file matches the uninterrupted run bit for bit
(positions, velocities, the interaction list and every slot array)."""
import pytest

from rfd.reflayout import UniaxialDriver

from test_formats import _resume_matches

pytestmark = pytest.mark.gpu


def test_checkpoint_resume_gpu(ctx, tmp_path):
    _resume_matches(lambda c: UniaxialDriver(ctx, c.R, c.mat, c.box, c.cmp, c.K),
                    lambda s: s.host(), tmp=tmp_path)
    assert ctx.flags() == 0


def test_lammps_export_of_native_system(ctx, tmp_path):
    """Positions, intact bonds and kinetic energies of a native system with a
    third of its bonds broken, read back from the LAMMPS file."""
    import numpy as np

    from rfd import formats, native_cases

    nc = native_cases.uniaxial_cylinder(ctx, 6.0, 8.0, seed=6)
    s = nc.sys
    N = s.N
    st, flag = s.bond_state()
    flag[::3] = 0  # a third of the bonds broken
    s.set_bond_state(st, flag)
    g = np.random.default_rng(2)
    s.set_velocities((g.standard_normal(3 * N) * 1e-3).astype(np.float32), None)
    s.step(5)
    intact = s.bond_counts()[0]
    p = tmp_path / "lammps.txt"
    formats.write_lammps_system(p, s)
    d = formats.read_lammps(p)
    st = s.state()
    R = st["R"].reshape(3, N).T
    np.testing.assert_allclose(d.R, R, rtol=1e-5, atol=1e-6)
    assert len(d.bonds) == intact and 0 < intact < s.bonds.n
    V = st["V"].reshape(3, N).T.astype(np.float64)
    m = s.params.rho * 4.0 / 3.0 * np.pi * s.radius.astype(np.float64) ** 3
    np.testing.assert_allclose(d.energy[:, 0], 0.5 * m * (V * V).sum(1), rtol=1e-5, atol=1e-30)
    assert ctx.flags() == 0
