"""The OpenMP build of the oracle (librfd_oracle_omp.so), the multi-core CPU
baseline: its per-particle loops are independent per particle, so a driver
run on it equals the single-thread build bit for bit (every slot array,
particle array and the loading history), here over 40 explicit steps of fast
uniaxial loading with contacts and breaking bonds."""
import numpy as np
import pytest

from oracle.oracle import OracleUniaxialDriver, lib
from rfd import cases


def test_openmp_build_matches_single_thread():
    threads = lib(omp=True).orc_threads()
    if threads < 2:
        pytest.skip("one OpenMP thread available")
    assert lib().orc_threads() == 1
    runs = []
    for omp in (False, True):
        c = cases.reference_uniaxial(8.0, 14.0, seed=21)
        c.cmp.v_load *= 150
        d = OracleUniaxialDriver(c.R, c.mat, c.box, c.cmp, c.grid(), c.K, omp=omp)
        d.setup()
        sums = [d.step() for _ in range(40)]
        runs.append((d, np.array(sums), c.cmp))
    (a, sa, ca), (b, sb, cb) = runs
    assert np.array_equal(sa, sb)
    for k in ("R", "V", "W", "F", "M", "IL", "ILtype", "iL"):
        assert np.array_equal(getattr(a.s, k), getattr(b.s, k)), k
    for k, v in a.s.slot.items():
        assert np.array_equal(v, b.s.slot[k]), k
    assert ca.stress == cb.stress and ca.stress > 0
