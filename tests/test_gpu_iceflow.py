"""The drop-in iceflow module: initialize/update/finalize on a State bag,
like the reference's tests/test_iceflow/test_iceflow.py, plus numeric checks
of its outputs against the oracle."""
import types

import numpy as np
import pytest
import torch

from igm_amd import config, iceflow
from igm_amd import synthetic as syn
from oracle import fields as OF
from oracle import vert_disc as VD

pytestmark = pytest.mark.gpu


class State:
    pass


def test_reference_smoke_zero_thickness(dev):
    """tests/test_iceflow/test_iceflow.py:5-24: 40 x 30 grid, thk = 0,
    usurf = 0, dX = 100; initialize/update/finalize; mean(ubar) < 100."""
    cfg = config.default_cfg()
    state = State()
    state.thk = np.zeros((30, 40), np.float32)
    state.usurf = np.zeros((30, 40), np.float32)
    state.dX = np.full((30, 40), 100.0, np.float32)
    state.it = -1
    iceflow.initialize(cfg, state)
    iceflow.update(cfg, state)
    iceflow.finalize(cfg, state)
    assert float(torch.mean(state.ubar)) < 100
    assert float(state.U.abs().max()) == 0.0  # masked: no ice anywhere


@pytest.mark.parametrize("method", ["solved", "emulated", "diagnostic"])
def test_methods_outputs(dev, method, tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)  # diagnostic writes errors_diagno.txt
    cfg = config.default_cfg()
    cfg.processes.iceflow.method = method
    cfg.processes.iceflow.emulator.fallback.check_its = 5
    g = syn.dome100()
    state = State()
    for k in ("thk", "usurf", "dX"):
        setattr(state, k, g[k])
    state.it = -1
    state.t = 0.0
    iceflow.initialize(cfg, state)
    state.it = 0
    iceflow.update(cfg, state)
    iceflow.finalize(cfg, state)
    nz = 10
    assert state.U.shape == (nz, 100, 100) and state.U.device.type == "cuda"
    for k in ("ubar", "vbar", "uvelsurf", "vvelsurf", "uvelbase", "vvelbase"):
        assert getattr(state, k).shape == (100, 100)
        assert torch.isfinite(getattr(state, k)).all()
    # derived fields = the oracle's get_vel* of the module's own U, V
    vw = VD.define_vertical_weight(nz, 4.0)
    out = OF.velocity_outputs(state.U.cpu().double(), state.V.cpu().double(), g["thk"], vw.astype(np.float64))
    for k in ("ubar", "vbar", "uvelsurf", "uvelbase"):
        assert torch.allclose(getattr(state, k).cpu().double(), out[k], rtol=1e-5, atol=1e-4)
    assert float(state.U.cpu()[:, g["thk"] <= 0].abs().max()) == 0.0
    np.testing.assert_allclose(state.vert_weight.ravel(), (6 * np.arange(10) + 13) / 400, atol=2e-7)
    if method == "solved":
        assert state.COST_Glen < 0  # gravity work dominates once the ice moves
        assert float(state.ubar.abs().max()) > 1.0


def test_unsupported_options_raise(dev):
    cfg = config.default_cfg()
    cfg.processes.iceflow.numerics.vert_basis = "Legendre"
    st = State()
    st.thk = np.ones((8, 8), np.float32); st.usurf = np.ones((8, 8), np.float32); st.dX = np.ones((8, 8), np.float32)
    with pytest.raises(NotImplementedError):
        iceflow.initialize(cfg, st)
    cfg = config.default_cfg()
    cfg.processes.iceflow.numerics.vert_basis = "Chebyshev"
    with pytest.raises(ValueError):
        iceflow.initialize(cfg, st)


def test_save_model_writes_a_keras_directory_that_loads_back(dev, tmp_path, monkeypatch):
    """finalize with save_model (emulate.py:331-359) writes iceflow-model/
    with model.h5 (Keras 2 HDF5), fieldin.dat and vert_grid.dat; pointing
    emulator.name at that directory loads the same weights."""
    monkeypatch.chdir(tmp_path)
    cfg = config.default_cfg()
    cfg.processes.iceflow.emulator.save_model = True
    g = syn.dome100()
    state = State()
    for k in ("thk", "usurf", "dX"):
        setattr(state, k, g[k])
    state.it = -1
    iceflow.initialize(cfg, state)
    iceflow.finalize(cfg, state)
    d = tmp_path / "iceflow-model"
    assert sorted(p.name for p in d.iterdir()) == ["fieldin.dat", "model.h5", "model.safetensors",
                                                   "vert_grid.dat"]
    w0 = state.iceflow_model.get_weights()
    cfg2 = config.default_cfg()
    cfg2.processes.iceflow.emulator.name = str(d)
    cfg2.processes.iceflow.emulator.retrain_freq = 0  # no retraining step at initialize
    state2 = State()
    for k in ("thk", "usurf", "dX"):
        setattr(state2, k, g[k])
    state2.it = -1
    iceflow.initialize(cfg2, state2)
    w1 = state2.iceflow_model.get_weights()
    assert len(w0) == len(w1)
    for (k0, b0), (k1, b1) in zip(w0, w1):
        assert np.array_equal(k0, k1) and np.array_equal(b0, b1)
