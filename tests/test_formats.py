"""The reference's on-disk formats (rfd/formats.py).

Each writer is checked line by line against the layout its reference writer
produces (header lines, field order, units: metres and pascals, C++'s
default 6-significant-digit output) on a hand-made case, and each reader
round-trips the writer's output.  The binary checkpoint is checked byte for
byte: 8 float3 of sample data, a 4-byte count (uint_fast32_t under MSVC), the
five float3 particle arrays and the float4 array, then the interaction list.
"""
import numpy as np
import pytest

from rfd import formats


def test_sample_txt(tmp_path):
    R = np.array([[1.0, 2.5, -3.0], [12.345678, 0.0, 100.0]])
    p = tmp_path / "Sample_0.txt"
    formats.write_sample(p, R)
    assert p.read_text().splitlines() == ["2", "0.001 0.0025 -0.003", "0.0123457 0 0.1"]
    back = formats.read_sample(p)
    np.testing.assert_allclose(back, [[1.0, 2.5, -3.0], [12.3457, 0.0, 100.0]], rtol=1e-6)


def test_stress_strain_txt(tmp_path):
    hist = np.array([[0.0, -0.001, 0.002], [-1e-4, -0.05, 0.0525]])
    p = tmp_path / "Uniaxial.txt"
    formats.write_stress_strain(p, hist)
    assert p.read_text().splitlines() == ["EpsZ StressZ_bottom StressZ_top",
                                          "0 -1e+06 2e+06", "-0.0001 -5e+07 5.25e+07"]
    np.testing.assert_allclose(formats.read_stress_strain(p), hist, rtol=1e-12)


def test_lammps_data(tmp_path):
    R = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    V = np.array([[1.0, 0, 0], [0, 0, 0], [0, 2.0, 0]])
    W = np.array([[0, 0, 3.0], [0, 0, 0], [0, 0, 0]])
    p = tmp_path / "lammps.txt"
    formats.write_lammps(p, R, [(0, 1), (0, 2)], (-1, -1, -1), (2, 2, 2), 0.03, V, W,
                         mass=2.0, inertia=0.5)
    lines = p.read_text().splitlines()
    assert lines[0] == "LAMMPS Description T0=0.03  (1st line of file)"
    assert lines[2:9] == ["3 atoms", "2 bonds", "0 angles", "0 dihedrals", "0 impropers",
                          "1 atom types", "2 bond types"]
    assert lines[9] == "-0.001 0.002 xlo xhi"
    k = lines.index("Atoms")
    assert lines[k + 2:k + 5] == ["1 1 0 0 0", "2 1 0.001 0 0", "3 1 0 0.001 0"]
    k = lines.index("Velocities")
    # 0.5 m v^2 and 0.5 I w^2, energy unit 1 J
    assert lines[k + 2:k + 5] == ["1 1 2.25 3.25", "2 0 0 0", "3 4 0 4"]
    k = lines.index("Bonds")
    assert lines[k + 2:k + 4] == ["1 2 1 2", "2 2 1 3"]
    d = formats.read_lammps(p)
    np.testing.assert_allclose(d.R, R, atol=1e-12)
    np.testing.assert_array_equal(d.bonds, [[0, 1], [0, 2]])
    np.testing.assert_allclose(d.energy[2], [4, 0, 4])


def test_checkpoint_binary_layout(tmp_path):
    g = np.random.default_rng(2)
    N, K = 4, 3
    S = N * K
    c = formats.Checkpoint(
        sample={k: g.standard_normal(3) for k in formats.SAMPLE_FIELDS},
        F=g.standard_normal(3 * N), V=g.standard_normal(3 * N), R=g.standard_normal(3 * N),
        M=g.standard_normal(3 * N), W=g.standard_normal(3 * N), Q=g.standard_normal(4 * N),
        IonP=K, acut=1.45, IL=g.integers(0, N, S).astype(np.uint32),
        ILtype=g.integers(0, 2, S).astype(np.uint8), iL=g.standard_normal(S),
        slot={k: g.standard_normal((S, 3)) for k in formats.SLOT_VEC})
    p = tmp_path / "DATA.dat"
    formats.write_checkpoint(p, c)
    b = p.read_bytes()
    assert len(b) == 8 * 12 + 4 + 5 * 12 * N + 16 * N + 4 + 4 + 4 + S * (4 + 1 + 4 + 7 * 12)
    # the count follows the sample data, the particle arrays follow the count
    assert np.frombuffer(b, np.uint32, 1, 96)[0] == N
    np.testing.assert_array_equal(np.frombuffer(b, np.float32, 3 * N, 100),
                                  c.F.astype(np.float32))
    il0 = 100 + 5 * 12 * N + 16 * N
    assert tuple(np.frombuffer(b, np.uint32, 2, il0)) == (S, K)
    assert np.frombuffer(b, np.float32, 1, il0 + 8)[0] == np.float32(1.45)
    d = formats.read_checkpoint(p)
    for k in ("F", "V", "R", "M", "W", "Q", "iL"):
        np.testing.assert_array_equal(getattr(d, k), np.asarray(getattr(c, k), np.float32))
    np.testing.assert_array_equal(d.IL, c.IL)
    np.testing.assert_array_equal(d.ILtype, c.ILtype)
    for k in formats.SLOT_VEC:
        np.testing.assert_array_equal(d.slot[k], c.slot[k].astype(np.float32))
    assert d.IonP == K and d.acut == pytest.approx(1.45)


def _resume_matches(make_driver, host, steps_before=15, steps_after=15, tmp=None):
    """Run a driver, checkpoint it through DATA.dat, restore into a fresh
    driver with a copy of the loading state, and continue both."""
    import copy

    from rfd import cases

    ca = cases.reference_uniaxial(6.0, 10.0, seed=3)
    ca.cmp.v_load *= 50
    a = make_driver(ca)
    a.setup()
    for _ in range(steps_before):
        a.step()
    p = tmp / "DATA.dat"
    formats.write_checkpoint(p, formats.checkpoint_from_layout(a.s, acut=ca.mat.a_farcut))
    cb = cases.reference_uniaxial(6.0, 10.0, seed=3)
    b = make_driver(cb)
    b.cmp = copy.deepcopy(a.cmp)
    formats.restore_layout(b.s, formats.read_checkpoint(p))
    for _ in range(steps_after):
        sa, sb = a.step(), b.step()
        assert np.array_equal(np.asarray(sa), np.asarray(sb))
    ha, hb = host(a.s), host(b.s)
    for k in ("R", "V", "W", "IL", "ILtype", "iL") + formats.SLOT_VEC:
        assert np.array_equal(np.asarray(ha[k]).view(np.uint8), np.asarray(hb[k]).view(np.uint8)), k
    assert a.cmp.strain == b.cmp.strain and a.cmp.stress == b.cmp.stress


def _oracle_host(s):
    return {k: getattr(s, k) for k in ("R", "V", "W", "IL", "ILtype", "iL")} | s.slot


def test_checkpoint_resume_oracle(tmp_path):
    """DATA.dat holds everything the per-step loop carries: a run continued
    from the file matches the uninterrupted run bit for bit (CPU oracle)."""
    from oracle.oracle import OracleUniaxialDriver

    _resume_matches(lambda c: OracleUniaxialDriver(c.R, c.mat, c.box, c.cmp, c.grid(), c.K),
                    _oracle_host, tmp=tmp_path)


def test_restore_rejects_other_sizes(tmp_path):
    from oracle.oracle import OracleLayout
    from rfd import cases

    c = cases.reference_uniaxial(6.0, 10.0, seed=3)
    ck = formats.checkpoint_from_layout(OracleLayout(c.R, c.K, c.grid()))
    with pytest.raises(ValueError):
        formats.restore_layout(OracleLayout(c.R[: 3 * 10], c.K, c.grid()), ck)


def test_particles_csv(tmp_path):
    """Metres about the sample centre; the real bounds of the read set."""
    p = tmp_path / "particles.csv"
    p.write_text("0.001 -0.002 0.0005\n-0.003,0.004,0.0\n0.0 0.0 0.01\n9 9 9\n")
    R, box = formats.read_particles_csv(p, 3, center=(10.0, 20.0, 30.0))
    np.testing.assert_allclose(R, [[11, 18, 30.5], [7, 24, 30], [10, 20, 40]], rtol=1e-6)
    np.testing.assert_allclose(box["A_real"], [7, 18, 30], rtol=1e-6)
    np.testing.assert_allclose(box["B_real"], [11, 24, 40], rtol=1e-6)
    np.testing.assert_allclose(box["size_real"], [4, 6, 10], rtol=1e-6)
    np.testing.assert_allclose(box["center_real"], [9, 21, 35], rtol=1e-6)
    with pytest.raises(ValueError):
        formats.read_particles_csv(p, 5)
