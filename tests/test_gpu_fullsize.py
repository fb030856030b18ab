"""The 1M-agent mixed crowd at its full size. The class-major scan of its 3M (class slot,
cell) counts spans 184 tiles at cell width 0.5 x range, so each tile's look-back walks
several 64-tile windows; at cell width 1.0 x range the 750k counts fit 46 tiles (one
window). The grid does not enter the results (neighbour lists are canonical), so both
widths give the same state bit for bit, over steps that also count the next step's
agents in the step kernels. A third run makes every scan tile give up its (bounded)
wait at once and sum its predecessors' counts itself: the same state again."""
import os

import numpy as np
import pytest

from umans_amd import gpu
from umans_amd import scenarios as S

pytestmark = pytest.mark.gpu


def test_full_size_mixed_crowd_independent_of_grid():
    sc = S.mixed()
    assert sc.n == 1_000_000
    outs = []
    for factor, spin in ((0.5, None), (1.0, None), (0.5, "-1")):
        if spin is not None:
            os.environ["UMANS_SCAN_SPIN_LIMIT"] = spin     # read when a context is created
        try:
            w = gpu.World(sc.world_type, sc.width, sc.height, sc.dt, sc.policies, sc.max_neighbors)
        finally:
            os.environ.pop("UMANS_SCAN_SPIN_LIMIT", None)
        w.upload(sc.agents)
        w.set_tuning(factor, 0, 0)
        w.step(3)
        outs.append(w.download())
        st = w.stats()
        w.close()
        assert st["steps"] == 3
    for k in ("pos", "vel", "viewdir"):
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)
        np.testing.assert_array_equal(outs[0][k], outs[2][k], err_msg=k)
    assert np.isfinite(outs[0]["pos"]).all()
