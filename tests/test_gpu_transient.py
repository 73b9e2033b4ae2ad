"""Transient evolution loop (smb_simple -> iceflow -> time -> thk) on the GPU
processes, small grid: runs, conserves sign and finiteness, and the solved
and emulated runs stay close in ice volume over a short period."""
import numpy as np
import pytest
import torch

from igm_amd import synthetic as syn
from igm_amd import transient as TR

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["solved", "emulated"])
def test_short_transient_run(dev, method):
    g = syn.dome100()
    cfg = TR.transient_cfg(method=method, years=5.0, retrain_freq=10)
    cfg.processes.iceflow.solver.nbitmax = 100
    st = TR.make_state(g)
    rec = TR.run(cfg, st)
    assert np.all(np.isfinite(rec["volume"])) and rec["t"][-1] == pytest.approx(5.0)
    assert float(st.thk.min()) >= 0.0
    assert np.all(rec["dt"] <= 1.0 + 1e-6)
    v0 = float(np.sum(g["thk"], dtype=np.float64)) * 100.0 ** 2 / 1e9
    # SMB = 0.005 (usurf - 2800) capped at 2: the volume changes smoothly
    assert abs(rec["volume"][-1] - v0) < 0.5 * v0
