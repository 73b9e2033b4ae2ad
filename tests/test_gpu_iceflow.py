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


def test_refine_polishes_with_its_schedule_and_adam_betas(dev):
    """iceflow.refine_emulated (emulator.refine.*) is the plain solver
    warm-started from the emulated field: the same field, bit for bit, as
    solve() with the schedule and Adam betas the keys name, masked to the ice;
    and it lowers the energy of the emulated field."""
    from igm_amd import solver as slv
    cfg = config.default_cfg()
    ice = cfg.processes.iceflow
    ice.method = "emulated"
    ice.emulator.retrain_freq = 0
    g = syn.dome100()
    state = State()
    for k in ("thk", "usurf", "dX"):
        setattr(state, k, g[k])
    state.it = -1
    iceflow.initialize(cfg, state)
    U0, V0 = state.U.clone(), state.V.clone()
    rf = ice.emulator.refine
    rf.its, rf.lr, rf.flat_frac, rf.lr_end, rf.beta1, rf.beta2 = 30, 3.0, 0.1, 0.3, 0.95, 0.99
    iceflow.refine_emulated(cfg, state)
    op = iceflow.energy_operator(cfg, state)
    r = slv.solve(op, U0, V0, 30, slv.LegacyAdamState(lr=3.0, beta1=0.95, beta2=0.99), early_stopping=False,
                  lr_schedule=slv.adam_schedule(3.0, 3, 27, 0.3))
    ice_cells = state.thk > 0
    for got, want in ((state.U, r.U), (state.V, r.V)):
        assert torch.equal(got, torch.where(ice_cells, want, torch.zeros_like(want)))
    # the Keras defaults give a different field (the betas are used)
    r_def = slv.solve(op, U0, V0, 30, slv.LegacyAdamState(lr=3.0), early_stopping=False,
                      lr_schedule=slv.adam_schedule(3.0, 3, 27, 0.3))
    assert not torch.equal(r_def.U, r.U)
    assert state.COST_REFINED < float(op.energy(U0, V0)[0, 3])
    # with a warm-up: 6 iterations from step 1 up to 3, 3 at 3, 21 decaying
    state.U, state.V = U0.clone(), V0.clone()
    rf.warmup_frac, rf.warmup_lr = 0.2, 1.0
    iceflow.refine_emulated(cfg, state)
    r_w = slv.solve(op, U0, V0, 30, slv.LegacyAdamState(lr=3.0, beta1=0.95, beta2=0.99), early_stopping=False,
                    lr_schedule=slv.adam_schedule(3.0, 3, 21, 0.3, n_warmup=6, lr_start=1.0))
    for got, want in ((state.U, r_w.U), (state.V, r_w.V)):
        assert torch.equal(got, torch.where(ice_cells, want, torch.zeros_like(want)))
    assert not torch.equal(r_w.U, r.U)


def test_energy_operator_cache_follows_its_inputs(dev):
    """iceflow.energy_operator reuses the float32 operator while the fields
    are the same tensors at the same version, and builds a new one after an
    in-place write or a rebinding of a field, or another physics value; a
    reused operator gives the energies of a freshly built one."""
    from igm_amd.energy import EnergyOperator
    cfg = config.default_cfg()
    g = syn.dome100()
    state = State()
    for k in ("thk", "usurf", "dX"):
        setattr(state, k, g[k])
    state.it = -1
    iceflow.initialize(cfg, state)
    U, V = state.U.contiguous(), state.V.contiguous()

    def fresh():
        fields = {k: getattr(state, k) for k in iceflow.FIELDIN}
        return EnergyOperator(np.asarray(state.dzeta), cfg.processes.iceflow.physics, fields=fields)

    op1 = iceflow.energy_operator(cfg, state)
    assert iceflow.energy_operator(cfg, state) is op1
    assert torch.equal(op1.energy(U, V), fresh().energy(U, V))
    state.thk.mul_(1.1)  # in place: the version changes
    op2 = iceflow.energy_operator(cfg, state)
    assert op2 is not op1
    assert torch.equal(op2.energy(U, V), fresh().energy(U, V))
    assert not torch.equal(op2.energy(U, V), op1.energy(U, V))
    state.usurf = state.usurf + 5.0  # rebound
    op3 = iceflow.energy_operator(cfg, state)
    assert op3 is not op2
    assert torch.equal(op3.energy(U, V), fresh().energy(U, V))
    cfg.processes.iceflow.physics.exp_glen = 4.0
    op4 = iceflow.energy_operator(cfg, state)
    assert op4 is not op3
    assert torch.equal(op4.energy(U, V), fresh().energy(U, V))
    assert iceflow.energy_operator(cfg, state) is op4


def test_fallback_replaces_a_failed_emulated_field(dev):
    """Energy-gap fallback: an emulator with zeroed weights predicts U = V = 0,
    whose energy lies far above what a few solver iterations reach, so the
    check falls back: fallback_count goes up by one per update and the
    module's U, V are the solver's field (the check's iterations, then the
    fallback solve with EarlyStopping, on the module's own optimizer),
    masked to the ice.  A second update with the same zero field falls back
    again."""
    from igm_amd import solver as slv
    cfg = config.default_cfg()
    ice = cfg.processes.iceflow
    ice.method = "emulated"
    ice.emulator.retrain_freq = 0
    fb = ice.emulator.fallback
    fb.check_its, fb.max_its = 10, 300
    g = syn.dome100()
    state = State()
    for k in ("thk", "usurf", "dX"):
        setattr(state, k, g[k])
    state.it = -1
    iceflow.initialize(cfg, state)
    model = state.iceflow_model
    model.set_weights([(np.zeros_like(k), np.zeros_like(b)) for k, b in model.get_weights()])
    for call in range(2):
        state.it = call
        t0 = state.optimizer.iterations
        iceflow.update(cfg, state)
        assert state.fallback_count == call + 1
        assert float(state.COST_EMULATED) > state.COST_CHECK + 0.01 * abs(state.COST_CHECK)
        # the same two solves on a copy of the optimizer at the same step
        op = iceflow.energy_operator(cfg, state)
        z = torch.zeros_like(state.U)
        opt = slv.LegacyAdamState(lr=float(ice.solver.step_size))
        opt.iterations = t0
        r1 = slv.solve(op, z, z.clone(), 10, opt, early_stopping=False)
        r2 = slv.solve(op, r1.U, r1.V, 300, opt, early_stopping=True)
        assert opt.iterations == state.optimizer.iterations
        ice_cells = state.thk > 0
        for got, want in ((state.U, r2.U), (state.V, r2.V)):
            assert torch.equal(got, torch.where(ice_cells, want, torch.zeros_like(want)))
        assert float(state.ubar.abs().max()) > 1.0


def test_fallback_keeps_a_good_emulated_field(dev):
    """The same check on a field that is already the solver's minimiser
    (here: the polished emulated field, then the check) does not fall back."""
    cfg = config.default_cfg()
    ice = cfg.processes.iceflow
    ice.method = "emulated"
    ice.emulator.retrain_freq = 0
    rf = ice.emulator.refine
    rf.its, rf.lr, rf.flat_frac, rf.lr_end, rf.beta1, rf.beta2 = 600, 30.0, 0.3, 0.01, 0.95, 0.99
    g = syn.dome100()
    state = State()
    for k in ("thk", "usurf", "dX"):
        setattr(state, k, g[k])
    state.it = -1
    iceflow.initialize(cfg, state)
    state.it = 0
    state.fallback_count = 0
    iceflow.update(cfg, state)
    assert state.fallback_count == 0
    gap = (float(state.COST_EMULATED) - state.COST_CHECK) / abs(state.COST_CHECK)
    assert gap <= 0.01


# The polish that the 1000 x 1000 benchmark chooses on its held-out fields
# (domes1000 seeds 21 and 22): 98 iterations, a geometric warm-up over the
# first 10 % from 3.33 to 20, 40 % at 20, then a geometric decay to 10; Adam
# (0.93, 0.91), epsilon 1e-9.
POLISH_DOMES1000 = {"its": 98, "warmup_frac": 0.1, "warmup_lr": 20.0 / 6.0, "lr": 20.0, "flat_frac": 0.4,
                    "lr_end": 10.0, "beta1": 0.93, "beta2": 0.91, "epsilon": 1e-9}


def test_benchmark_polish_generalises_to_an_unseen_field(dev):
    """The emulated update (this build's trained network, that polish, the
    energy-gap check) on domes1000 seed 31 -- a field that neither the
    network's training nor the schedule's choice saw -- is within 1 % of
    the field's float64 minimiser (28,000 float32 + 2,000 float64
    iterations) and passes the check."""
    from igm_amd.energy import EnergyOperator
    from igm_amd.solver import converged_solution, rel_err
    print("case domes1000 seed 31: emulated update with the benchmark's polish", flush=True)
    cfg = config.default_cfg()
    ice = cfg.processes.iceflow
    ice.method = "emulated"
    ice.emulator.retrain_freq = 0
    for k, v in POLISH_DOMES1000.items():
        setattr(ice.emulator.refine, k, v)
    from igm_amd import emulator as EM
    ice.emulator.pretrained = True
    ice.emulator.name = EM.os.path.join(EM.EMULATOR_DIR, "igm_amd_cnn_16_32_nz10.safetensors")
    g = syn.domes1000(seed=31)
    state = State()
    for k in ("thk", "usurf", "arrhenius", "slidingco", "dX"):
        setattr(state, k, g[k])
    state.it = -1
    iceflow.initialize(cfg, state)
    state.it = 0
    iceflow.update(cfg, state)
    assert state.fallback_count == 0
    gap = (float(state.COST_EMULATED) - state.COST_CHECK) / abs(state.COST_CHECK)
    f32 = {k: torch.as_tensor(g[k], device=dev) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")}
    dz = np.asarray(state.dzeta)
    op32 = EnergyOperator(dz, ice.physics, fields=f32)
    op64 = EnergyOperator(dz, ice.physics, fields={k: v.double() for k, v in f32.items()}, dtype=torch.float64)
    z = torch.zeros_like(state.U)
    Ur, Vr, _, _ = converged_solution(op32, op64, z, z.clone())
    err = rel_err(state.U, state.V, Ur, Vr, f32["thk"] > 0)
    print(f"  rel err {err:.4f}, energy gap at the check {gap:.2e}", flush=True)
    assert err <= 0.01
