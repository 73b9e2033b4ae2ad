"""Vertical velocity W from U, V (CPU restatement, numpy float32).

Test infrastructure only: the checker of the GPU kernel, never called by
the product path.

Restates the default of ``igm/processes/vert_flow/vert_flow.py:21-56``
(Lagrange basis, version 2, method ``kinematic``):
``compute_vertical_velocity_kinematic_v2`` / ``compute_w_kinematic_tf``
(``vert_flow_v2.py:17-37``) with ``compute_gradient_layers_tf`` (``61-87``)
and ``compute_divflux_layers`` (``125-147``), in the reference's op order.

    z_k   = thk * levels[k]                       (cumsum of compute_dz, levels[0] = 0)
    s_k   = topg + z_k                            (elevation of level k)
    ub_k  = sum_{l<=k} w_l U_l / sum_{l<=k} w_l   (w = vert_weight)
    W_k   = -div(ub_k z_k, vb_k z_k) + U_k ds_k/dx + V_k ds_k/dy

with the reference's stencils: the slope of s is centred inside and
one-sided at the edges (edge values extrapolated 1.5 s_0 - 0.5 s_1), the
flux divergence is centred inside and half one-sided at the edges.

Pins (``tests/test_oracle_vert_flow.py``): a uniform flow over a flat bed
with uniform thickness gives W = 0; a flow u = a x over a flat bed gives the
incompressible W(z) = -a z inside; a uniform flow over a bed of slope b gives
W = u b at every level (the flow follows the bed); edge stencils against
hand-computed values.
"""
import numpy as np

f32 = np.float32


def gradient_layers(s, dx, dy):
    """vert_flow_v2.py:61-87 on (Nz, Ny, Nx): centred inside, one-sided edges."""
    s = s.astype(f32)
    EX = np.concatenate([f32(1.5) * s[:, :, 0:1] - f32(0.5) * s[:, :, 1:2],
                         f32(0.5) * s[:, :, :-1] + f32(0.5) * s[:, :, 1:],
                         f32(1.5) * s[:, :, -1:] - f32(0.5) * s[:, :, -2:-1]], axis=2)
    diffx = (EX[:, :, 1:] - EX[:, :, :-1]) / f32(dx)
    EY = np.concatenate([f32(1.5) * s[:, 0:1, :] - f32(0.5) * s[:, 1:2, :],
                         f32(0.5) * s[:, :-1, :] + f32(0.5) * s[:, 1:, :],
                         f32(1.5) * s[:, -1:, :] - f32(0.5) * s[:, -2:-1, :]], axis=1)
    diffy = (EY[:, 1:, :] - EY[:, :-1, :]) / f32(dy)
    return diffx.astype(f32), diffy.astype(f32)


def divflux_layers(u, v, h, dx, dy):
    """vert_flow_v2.py:125-147: centred divergence of (u h, v h) per layer."""
    Qx = (u * h).astype(f32)
    Qy = (v * h).astype(f32)
    dQx = np.concatenate([Qx[:, :, 0:1], f32(0.5) * (Qx[:, :, :-1] + Qx[:, :, 1:]), Qx[:, :, -1:]], axis=2)
    dQy = np.concatenate([Qy[:, 0:1, :], f32(0.5) * (Qy[:, :-1, :] + Qy[:, 1:, :]), Qy[:, -1:, :]], axis=1)
    return ((dQx[:, :, 1:] - dQx[:, :, :-1]) / f32(dx) + (dQy[:, 1:, :] - dQy[:, :-1, :]) / f32(dy)).astype(f32)


def w_kinematic(U, V, topg, thk, levels, vert_weight, dx):
    """vert_flow_v2.py:17-37: W (Nz, Ny, Nx) float32."""
    U = np.asarray(U, f32)
    V = np.asarray(V, f32)
    thk = np.asarray(thk, f32)
    levels = np.asarray(levels, f32)
    w = np.asarray(vert_weight, f32).reshape(-1, 1, 1)
    ddz = levels[1:] - levels[:-1]
    dz = thk[None] * ddz[:, None, None]                                    # compute_dz
    cumdz = np.concatenate([np.zeros_like(thk)[None], np.cumsum(dz, axis=0, dtype=f32)], axis=0)
    sx, sy = gradient_layers(np.asarray(topg, f32)[None] + cumdz, dx, dx)
    ub = np.cumsum(w * U, axis=0, dtype=f32) / np.cumsum(w * np.ones_like(U), axis=0, dtype=f32)
    vb = np.cumsum(w * V, axis=0, dtype=f32) / np.cumsum(w * np.ones_like(V), axis=0, dtype=f32)
    div = divflux_layers(ub, vb, cumdz, dx, dx)
    return (-div + U * sx + V * sy).astype(f32)
