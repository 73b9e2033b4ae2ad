"""Vertical discretisation, Lagrange basis (host-side setup, float32).

igm/processes/iceflow/vert_disc.py:10-47: stretched levels
levels_k = (zeta_k / vs) (1 + (vs - 1) zeta_k) with zeta_k = k/(Nz-1),
layer thicknesses dzeta = diff(levels), midpoints zeta, and the velbar
weights built on Nz+1 points.  These are a few dozen numbers computed once.
"""
import numpy as np


def compute_levels(Nz, vert_spacing):
    zeta = (np.arange(Nz, dtype=np.float64) / (Nz - 1)).astype(np.float32)
    vs = np.float32(vert_spacing)
    return (zeta / vs) * (np.float32(1) + (vs - np.float32(1)) * zeta)


def compute_zeta_dzeta(levels):
    levels = np.asarray(levels, np.float32)
    if levels.shape[0] > 1:
        return (levels[1:] + levels[:-1]) / np.float32(2), levels[1:] - levels[:-1]
    return np.full(1, 0.5, np.float32), np.ones(1, np.float32)


def define_vertical_weight(Nz, vert_spacing):
    zeta = (np.arange(Nz + 1, dtype=np.float64) / Nz).astype(np.float32)
    vs = np.float32(vert_spacing)
    w = (zeta / vs) * (np.float32(1) + (vs - np.float32(1)) * zeta)
    return (w[1:] - w[:-1])[:, None, None]
