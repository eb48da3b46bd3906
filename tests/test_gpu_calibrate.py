"""The calibration outer loop on the GPU with small test specimens: one CG
step evaluates 6 parameter sets x 2 tests in batches; every peak is finite
and positive, the bond-radius variation raises the uniaxial peak (a 5 %
thicker bond is a stronger bond), and the step moves the parameters."""
import numpy as np
import pytest

from rfd import calibrate as cal
from rfd.params import VELOCITY

pytestmark = pytest.mark.gpu


def test_one_cg_step(ctx):
    geom = cal.TestGeometry(uni_D=8.0, uni_H=16.0, bra_D=16.0, bra_T=6.0,
                            uni_v=-5 * 10.0 * 0.02 / VELOCITY)
    ev = cal.Evaluator(ctx, geom, max_steps=40000, chunk=2000)
    st = cal.CGState(params=cal.GRANITE_START.copy(), goals=cal.GRANITE_GOALS.copy())
    sets = st.parameter_sets()
    R = ev.evaluate(sets)
    print("results", R.tolist(), flush=True)
    assert np.all(np.isfinite(R)) and np.all(R > 0)
    assert R[4, 1] > R[0, 1]
    p0 = st.params.copy()
    st.update(R)
    assert np.all(np.isfinite(st.params)) and not np.allclose(st.params, p0)
    assert ctx.flags() == 0
