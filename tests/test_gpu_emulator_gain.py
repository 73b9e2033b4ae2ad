"""What the trained emulator gains over the plain method.

On a held-out 400 x 400 multi-dome field (dx = 50 m, A uniform, seed 21: not
a training, validation or benchmark seed), the plain solver warm-started from
the trained network's field is compared with the same solver started from
zero, at equal iteration budgets of the same schedule family (lr 30 for 30 %
of the budget, then a geometric decay to 1e-3), against the field's
converged float64 minimiser.  The warm start must be ahead at every budget
and reach 1 % in fewer iterations.
"""
import os

import numpy as np
import pytest
import torch

from igm_amd import emulator as EM
from igm_amd import synthetic as syn
from igm_amd.energy import EnergyOperator
from igm_amd.solver import LegacyAdamState, adam_schedule, converged_solution, rel_err, solve
from oracle.energy import DEFAULT_PHYSICS
from tests.helpers import dzeta_for

pytestmark = pytest.mark.gpu

KEYS = ("thk", "usurf", "arrhenius", "slidingco", "dX")
WEIGHTS = os.path.join(EM.EMULATOR_DIR, "igm_amd_cnn_16_32_nz10.safetensors")


def _polish(op, U0, V0, n):
    nf = int(0.3 * n)
    r = solve(op, U0, V0, n, LegacyAdamState(lr=30.0), early_stopping=False,
              lr_schedule=adam_schedule(30.0, nf, n - nf))
    return r.U, r.V


def test_trained_emulator_beats_cold_start(dev):
    nz, shape, dx, seed = 10, (400, 400), 50.0, 21
    g = syn._pack(syn._bed(shape, dx, seed=seed), syn.random_domes(shape, dx, 6, seed=seed),
                  syn.DEFAULT_A, syn.DEFAULT_C, dx)
    dz = dzeta_for(nz)
    f32 = {k: torch.as_tensor(g[k], device=dev) for k in KEYS}
    f64 = {k: v.double() for k, v in f32.items()}
    op32 = EnergyOperator(dz, DEFAULT_PHYSICS, fields=f32)
    op64 = EnergyOperator(dz, DEFAULT_PHYSICS, fields=f64, dtype=torch.float64)
    z = torch.zeros((nz,) + shape, device=dev)
    Ur, Vr, _, change = converged_solution(op32, op64, z, z.clone())
    assert change < 1e-5
    mask = f32["thk"] > 0

    weights, _ = EM.load_weights(WEIGHTS)
    cnn = EM.CNN(nz, 16, device=dev, weights=weights)
    keep = [f32[k] for k in KEYS]
    Y = cnn.forward(EM.planar_fields(keep))
    Ue = Y[0, ..., :nz].permute(2, 0, 1).contiguous()
    Ve = Y[0, ..., nz:].permute(2, 0, 1).contiguous()
    e_emu = rel_err(Ue, Ve, Ur, Vr, mask)
    print(f"emulator alone: rel err {e_emu:.4f}", flush=True)
    assert e_emu < 0.5

    first_warm = first_cold = None
    for n in (250, 500, 1000, 2000, 4000):
        ew = rel_err(*_polish(op32, Ue, Ve, n), Ur, Vr, mask)
        ec = rel_err(*_polish(op32, z, z.clone(), n), Ur, Vr, mask)
        print(f"{n} iterations: warm start {ew:.4f}, cold start {ec:.4f}", flush=True)
        assert ew < ec
        if first_warm is None and ew <= 0.01:
            first_warm = n
        if first_cold is None and ec <= 0.01:
            first_cold = n
    assert first_warm is not None
    assert first_cold is None or first_warm < first_cold
