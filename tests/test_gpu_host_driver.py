"""The C++ host driver (rockfit-dem_amd/host/rfd_run): the reference's test
loops written in C++ over the C ABI, as a reference maintainer's host code
would call it.

It reads the particles from a DATA.dat written by rfd.formats, builds the
lists and bonds, runs explicit steps of the uniaxial or Brazilian test, and
writes the stress-strain file and the final DATA.dat.  Both are compared with
the oracle's driver on the same particles and parameters: the final state bit
for bit (positions, velocities, spins, loads, the interaction list and the
slot arrays), the stress-strain record to the 6 significant digits of the
text format.
"""
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleUniaxialDriver
from rfd import cases, formats
from rfd.build import RUN

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", ["uniaxial", "brazil"])
def test_cpp_driver_matches_oracle(mode, tmp_path):
    assert RUN.exists(), "the C++ host driver is not built"
    mk = cases.reference_uniaxial if mode == "uniaxial" else cases.reference_brazil
    D, H = (8.0, 14.0) if mode == "uniaxial" else (14.0, 5.0)
    scale, steps = 50.0, 80
    c = mk(D, H, seed=2)
    N, K = c.N, c.K
    S = N * K
    z = np.zeros(3 * N, np.float32)
    slots = {k: np.zeros((S, 3), np.float32) for k in formats.SLOT_VEC}
    ck = formats.Checkpoint({k: np.zeros(3) for k in formats.SAMPLE_FIELDS}, z, z, c.R, z, z,
                            np.zeros(4 * N, np.float32), K, 0.0,
                            np.full(S, 0xFFFFFFFF, np.uint32), np.zeros(S, np.uint8),
                            np.zeros(S, np.float32), slots)
    formats.write_checkpoint(tmp_path / "in.dat", ck)
    out = tmp_path / "out"
    out.mkdir()
    print(f"{mode}: N={N}, {steps} steps", flush=True)
    r = subprocess.run([str(RUN), "--mode", mode, "--D", str(D), "--H", str(H), "--in",
                        str(tmp_path / "in.dat"), "--out", str(out), "--steps", str(steps),
                        "--vload-scale", str(scale)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "flags=0" in r.stdout

    c.cmp.v_load *= scale
    d = OracleUniaxialDriver(c.R, c.mat, c.box, c.cmp, c.grid(), c.K)
    d.setup()
    for _ in range(steps):
        d.step()
    got = formats.read_checkpoint(out / "DATA.dat")
    assert got.IonP == K and len(got.IL) == S
    for k in ("R", "V", "W", "F", "M", "IL", "ILtype", "iL"):
        a, b = np.asarray(getattr(got, k)), getattr(d.s, k)
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
    for k in formats.SLOT_VEC:
        assert np.array_equal(got.slot[k].reshape(-1), d.s.slot[k]), k
    hist = formats.read_stress_strain(out / ("Uniaxial.txt" if mode == "uniaxial" else
                                             "Brazil.txt"))
    ref = np.array(c.cmp.history)
    assert hist.shape == ref.shape == (steps, 3)
    assert np.abs(ref[:, 1:]).max() > 0  # the platens carry load
    np.testing.assert_allclose(hist, ref, rtol=1e-5, atol=1e-12)
