"""Host logic of the time process (no GPU: cfl = 0 needs no velocity max)."""
import numpy as np
import torch

from igm_amd import config
from igm_amd.processes import time as timemod


class _S:
    pass


def _run(start, end, save, step_max):
    cfg = config.default_cfg()
    c = cfg.processes.time
    c.start, c.end, c.save, c.cfl, c.step_max = start, end, save, 0.0, step_max
    st = _S()
    st.thk = torch.zeros((2, 2))
    st.dx = 100.0
    timemod.initialize(cfg, st)
    st.it = -1
    ts, saves = [], []
    st.continue_run = True
    while st.continue_run:
        timemod.update(cfg, st)
        ts.append(st.t)
        if st.saveresult:
            saves.append(st.t)
        st.it += 1
        assert len(ts) < 100000
    return ts, saves


def test_run_ends_exactly_on_the_end_time():
    # float32 steps that do not divide the interval: the last step is cut to
    # land on the save/end time, and the loop stops there (time.py:53-63)
    for step in (0.3, 0.0056, 0.7, 1.0 / 3.0):
        ts, saves = _run(0.0, 10.0, 10.0, step)
        assert ts[-1] == np.float32(10.0)
        assert saves[-1] == np.float32(10.0)
        assert all(isinstance(t, np.float32) for t in ts)


def test_save_times_are_hit():
    ts, saves = _run(2000.0, 2100.0, 10.0, 0.7)
    assert saves == [np.float32(x) for x in np.arange(2000.0, 2101.0, 10.0)]
