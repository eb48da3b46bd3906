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
