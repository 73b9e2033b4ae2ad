"""Pins of the remaining oracle pieces: vertical discretisation closed forms,
legacy Adam by hand, EarlyStopping sequences, CNN structure and Keras
cross-correlation semantics.  CPU only."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import cnn as OC
from oracle import fields as FL
from oracle import solver as S
from oracle import vert_disc as VD


def test_vertical_closed_forms_nz10_vs4():
    """levels_k = k(k+3)/108, dzeta_l = (l+2)/54, vert_weight_k = (6k+13)/400."""
    k = np.arange(10)
    lev = VD.compute_levels(10, 4.0)
    assert np.allclose(lev, k * (k + 3) / 108.0, rtol=0, atol=2e-7)
    zeta, dz = VD.compute_zeta_dzeta(lev)
    assert np.allclose(dz, (np.arange(9) + 2) / 54.0, atol=2e-7)
    assert abs(float(dz.sum()) - 1.0) < 1e-6
    assert np.allclose(zeta, (lev[1:] + lev[:-1]) / 2, atol=1e-7)
    vw = VD.define_vertical_weight(10, 4.0).ravel()
    assert np.allclose(vw, (6 * k + 13) / 400.0, atol=2e-7)
    assert abs(float(vw.sum()) - 1.0) < 1e-6


def test_legacy_adam_by_hand():
    """Two steps of ResourceApplyAdam on a scalar, computed by hand in float64."""
    x = torch.tensor([1.0], dtype=torch.float64)
    m = torch.zeros_like(x); v = torch.zeros_like(x)
    opt = S.LegacyAdam(lr=0.5)
    g1, g2 = 2.0, -1.0
    opt.apply([x], [torch.tensor([g1], dtype=torch.float64)], [m], [v])
    m1 = 0.1 * g1; v1 = 0.001 * g1 * g1
    a1 = 0.5 * np.sqrt(1 - 0.999) / (1 - 0.9)
    x1 = 1.0 - m1 * a1 / (np.sqrt(v1) + 1e-7)
    assert float(x) == pytest.approx(x1, rel=1e-14)
    opt.apply([x], [torch.tensor([g2], dtype=torch.float64)], [m], [v])
    m2 = m1 + (g2 - m1) * 0.1; v2 = v1 + (g2 * g2 - v1) * 0.001
    a2 = 0.5 * np.sqrt(1 - 0.999 ** 2) / (1 - 0.9 ** 2)
    assert float(x) == pytest.approx(x1 - m2 * a2 / (np.sqrt(v2) + 1e-7), rel=1e-14)


def test_legacy_adam_equals_torch_adam_without_epsilon_correction():
    """With epsilon -> 0 the legacy update equals torch.optim.Adam's (the two
    differ only in where epsilon enters)."""
    x1 = torch.tensor([0.3, -2.0], dtype=torch.float64)
    x2 = x1.clone().requires_grad_(True)
    m = torch.zeros_like(x1); v = torch.zeros_like(x1)
    opt = S.LegacyAdam(lr=0.1, epsilon=0.0)
    topt = torch.optim.Adam([x2], lr=0.1, eps=0.0)
    for step in range(5):
        g = torch.tensor([np.sin(step + 1.0), np.cos(step * 2.0)], dtype=torch.float64)
        opt.apply([x1], [g], [m], [v])
        x2.grad = g.clone(); topt.step()
    assert torch.allclose(x1, x2.detach(), rtol=1e-12, atol=1e-14)


def test_early_stopping_sequences():
    es = S.EarlyStopping(relative_min_delta=0.01, patience=3)
    seq = [10.0, 9.0, 8.95, 8.94, 8.93]
    out = [es.should_stop(c) for c in seq]
    assert out == [False, False, False, False, True]
    es = S.EarlyStopping(relative_min_delta=0.01, patience=2)
    out = [es.should_stop(c) for c in [-5.0, -6.0, -6.01, -7.0, -7.001, -7.002]]
    assert out == [False, False, False, False, False, True]


def test_cnn_parameter_count_and_keras_semantics():
    w = OC.init_weights(5, 20, 16, 32, 3)
    assert OC.count_params(w) == 140852
    # Keras "same" Conv2D is a cross-correlation with zero padding: a kernel
    # that is 1 at tap (r, s) reads pixel (y + r - 1, x + s - 1)
    X = torch.arange(1.0, 1.0 + 5 * 6, dtype=torch.float64).reshape(1, 5, 6, 1)
    k = np.zeros((3, 3, 1, 1)); k[2, 0, 0, 0] = 1.0
    Y = OC.forward(X, [(k, np.zeros(1)), (np.ones((1, 1, 1, 1)), np.zeros(1))], dtype=torch.float64)
    # hidden layer leaky_relu(identity of positive values) then 1x1 identity
    assert float(Y[0, 1, 2, 0]) == float(X[0, 2, 1, 0])
    assert float(Y[0, 4, 2, 0]) == 0.0  # row 5 is outside -> zero padding


def test_layout_round_trip():
    U = torch.randn(10, 4, 5, dtype=torch.float64); V = torch.randn(10, 4, 5, dtype=torch.float64)
    Y = FL.UV_to_Y(U, V)
    assert Y.shape == (1, 4, 5, 20)
    U2, V2 = FL.Y_to_UV(Y, 10)
    assert torch.equal(U2[0], U) and torch.equal(V2[0], V)


def test_velocity_outputs_mask_and_weights():
    nz = 10
    U = torch.ones(nz, 3, 3, dtype=torch.float64) * 5.0
    thk = np.ones((3, 3)); thk[1, 1] = 0.0
    out = FL.velocity_outputs(U, 2 * U, thk, VD.define_vertical_weight(nz, 4.0).astype(np.float64))
    assert float(out["U"][:, 1, 1].abs().sum()) == 0.0
    assert float(out["ubar"][0, 0]) == pytest.approx(5.0, rel=1e-6)   # weights sum to one
    assert float(out["vbar"][2, 2]) == pytest.approx(10.0, rel=1e-6)
