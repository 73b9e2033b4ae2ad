"""Vertical discretisation (Lagrange basis), float32 as in the reference.

Restates ``igm/processes/iceflow/vert_disc.py``.  Pinned by the closed forms
for Nz=10, vert_spacing=4 (levels_k = k(k+3)/108, dzeta_l = (l+2)/54,
vert_weight_k = (6k+13)/400) checked in ``tests/test_oracle_misc.py``.
"""
import numpy as np


def compute_levels(Nz, vert_spacing):
    """vert_disc.py:10-13 -- stretched levels in [0, 1], float32."""
    zeta = (np.arange(Nz, dtype=np.float64) / (Nz - 1)).astype(np.float32)
    vs = np.float32(vert_spacing)
    return (zeta / vs) * (np.float32(1.0) + (vs - np.float32(1.0)) * zeta)


def compute_zeta_dzeta(levels):
    """vert_disc.py:26-35 -- layer midpoints and thicknesses, float32."""
    levels = np.asarray(levels, dtype=np.float32)
    if levels.shape[0] > 1:
        zeta = (levels[1:] + levels[:-1]) / np.float32(2)
        dzeta = levels[1:] - levels[:-1]
        return zeta, dzeta
    return np.full(1, 0.5, np.float32), np.ones(1, np.float32)


def define_vertical_weight(Nz, vert_spacing):
    """vert_disc.py:43-47 -- weights used by get_velbar, shape (Nz, 1, 1).

    Built on Nz+1 points (not on the levels grid), as the reference does.
    """
    zeta = (np.arange(Nz + 1, dtype=np.float64) / Nz).astype(np.float32)
    vs = np.float32(vert_spacing)
    weight = (zeta / vs) * (np.float32(1.0) + (vs - np.float32(1.0)) * zeta)
    return (weight[1:] - weight[:-1])[:, None, None]
