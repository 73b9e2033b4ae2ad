"""Pins of the vert_flow oracle (kinematic W, version 2) by closed forms."""
import numpy as np

from oracle import vert_flow as VF
from oracle import vert_disc as VD


def _setup(nz=10, ny=12, nx=15):
    levels = np.asarray(VD.compute_levels(nz, 4.0), np.float64)
    w = np.asarray(VD.define_vertical_weight(nz, 4.0), np.float64).ravel()
    return levels, w


def test_uniform_flow_flat_bed_has_no_vertical_velocity():
    levels, w = _setup()
    nz, ny, nx = 10, 12, 15
    U = np.full((nz, ny, nx), 37.0)
    V = np.full((nz, ny, nx), -12.0)
    W = VF.w_kinematic(U, V, np.full((ny, nx), 2500.0), np.full((ny, nx), 300.0), levels, w, 50.0)
    assert np.abs(W).max() <= 1e-3


def test_extending_flow_is_incompressible_inside():
    # u = a x in every level, flat bed, uniform H: div(u z_k) = a z_k inside,
    # so W(z) = -a z (du/dx + dw/dz = 0)
    levels, w = _setup()
    nz, ny, nx, dx, a, H = 10, 12, 15, 50.0, 0.01, 300.0
    x = np.arange(nx) * dx
    U = np.broadcast_to(a * x, (nz, ny, nx)).copy()
    V = np.zeros((nz, ny, nx))
    W = VF.w_kinematic(U, V, np.zeros((ny, nx)), np.full((ny, nx), H), levels, w, dx)
    z = (H * levels)[:, None, None]
    np.testing.assert_allclose(W[:, :, 1:-1], np.broadcast_to(-a * z, (nz, ny, nx - 2)), rtol=1e-5, atol=1e-4)
    # edges: half one-sided divergence (Q_1 - Q_0) / (2 dx) = a z / 2
    np.testing.assert_allclose(W[:, :, 0], np.broadcast_to(-0.5 * a * z[:, :, 0], (nz, ny)), rtol=1e-5, atol=1e-4)


def test_flow_follows_an_inclined_bed():
    # uniform u over a bed of slope b and uniform H: no divergence, W = u b
    levels, w = _setup()
    nz, ny, nx, dx, b, u = 10, 12, 15, 50.0, -0.05, 80.0
    x = np.arange(nx) * dx
    topg = np.broadcast_to(2500.0 + b * x, (ny, nx))
    U = np.full((nz, ny, nx), u)
    V = np.zeros((nz, ny, nx))
    W = VF.w_kinematic(U, V, topg, np.full((ny, nx), 200.0), levels, w, dx)
    np.testing.assert_allclose(W, u * b, rtol=1e-4, atol=1e-3)


def test_edge_stencils_by_hand():
    s = np.array([[[1.0, 4.0, 9.0, 16.0]]], np.float32)
    gx, gy = VF.gradient_layers(np.concatenate([s, s], axis=1), 1.0, 1.0)
    # one-sided at the edges, centred inside
    np.testing.assert_allclose(gx[0, 0], [3.0, 4.0, 6.0, 7.0])
    np.testing.assert_allclose(gy, 0.0)
    Q = VF.divflux_layers(np.ones((1, 1, 4), np.float32), np.zeros((1, 1, 4), np.float32), s, 1.0, 1.0)
    np.testing.assert_allclose(Q[0, 0], [1.5, 4.0, 6.0, 3.5])
