"""Pins of the transport restatement (oracle/transport.py). CPU only."""
import numpy as np
import pytest

from oracle import transport as T


def test_limiters():
    a = np.array([1.0, -2.0, 3.0, 1.0, 0.0]); b = np.array([2.0, -1.0, -1.0, 1.0, 5.0])
    assert np.array_equal(T.minmod(a, b), [1.0, -1.0, 0.0, 0.0, 0.0])  # equal |a|=|b| -> 0
    assert np.array_equal(T.maxmod(a, b), [2.0, -2.0, 0.0, 0.0, 0.0])


def test_divflux_conservation():
    """Sum of div(q) dx dy = net boundary flux = 0 for ice away from the edges."""
    rng = np.random.default_rng(0)
    ny, nx, dx = 40, 50, 100.0
    h = np.zeros((ny, nx), np.float32); h[10:30, 10:40] = rng.uniform(50, 300, (20, 30))
    u = rng.normal(0, 50, (ny, nx)).astype(np.float32); v = rng.normal(0, 50, (ny, nx)).astype(np.float32)
    for slope in ("godunov", "minmod", "superbee"):
        d = T.compute_divflux_slope_limiter(u, v, h, dx, dx, 0.5, slope)
        assert abs(float(np.sum(d.astype(np.float64)))) <= 1e-4 * float(np.sum(np.abs(d)))


def test_godunov_is_first_order_upwind():
    ny, nx, dx = 3, 20, 100.0
    h = np.linspace(0, 100, nx, dtype=np.float32)[None].repeat(ny, 0); h[:, :3] = 0; h[:, -3:] = 0
    u = np.full((ny, nx), 10.0, np.float32); v = np.zeros((ny, nx), np.float32)
    d = T.compute_divflux_slope_limiter(u, v, h, dx, dx, 1.0, "godunov")
    q = 10.0 * h  # upwind flux through the east face of each cell with u > 0
    expect = (q - np.concatenate([np.zeros((ny, 1)), q[:, :-1]], 1)) / dx
    assert np.allclose(d, expect, rtol=1e-5, atol=1e-5)


def test_uniform_translation_superbee():
    """A smooth bump advected at constant u for a CFL-0.5 step moves right;
    superbee keeps it non-negative and conserves its mass."""
    nx, dx, u0, dt = 200, 100.0, 50.0, 1.0
    x = np.arange(nx) * dx
    h = (100 * np.exp(-((x - 5000) / 1000) ** 2)).astype(np.float32)[None].repeat(3, 0)
    thk = h.copy()
    for _ in range(40):
        thk, _, _, _ = T.thk_update(thk, np.full_like(thk, u0), np.zeros_like(thk), np.zeros_like(thk),
                                    np.zeros_like(thk), dx, dt, "superbee")
    assert thk.min() >= 0.0
    assert float(thk[1].sum()) == pytest.approx(float(h[1].sum()), rel=1e-5)
    com0 = float((h[1] * x).sum() / h[1].sum()); com = float((thk[1] * x).sum() / thk[1].sum())
    assert com - com0 == pytest.approx(u0 * dt * 40, rel=0.02)


def test_cfl_and_smb():
    assert T.cfl_dt(np.array([[0.0, -30.0]]), np.array([[10.0, 0.0]]), 50.0, 0.3, 1.0) == pytest.approx(0.5)
    assert T.cfl_dt(np.zeros((2, 2)), np.zeros((2, 2)), 50.0) == 1.0
    s = T.smb_simple(np.array([2700.0, 2800.0, 3500.0, 4000.0]), 0.005, 0.005, 2800.0, 2.0)
    assert np.allclose(s, [-0.5, 0.0, 2.0, 2.0])
