"""Shared test inputs: seeded glaciers and smooth velocity fields."""
import numpy as np
import torch

from igm_amd import synthetic as syn
from oracle import vert_disc as VD


def dzeta_for(nz, vs=4.0):
    return VD.compute_zeta_dzeta(VD.compute_levels(nz, vs))[1]


def small_glacier(ny, nx, dx=100.0, seed=0, A=None, C=None):
    """A dome on a rough inclined bed, sized to the grid (ny, nx)."""
    rng = np.random.default_rng(seed)
    bed = 2500.0 - 0.1 * np.arange(nx)[None, :] * dx + syn.grf((ny, nx), 20.0, 3 * dx, dx, seed)
    R = 0.45 * min(ny, nx) * dx
    H = syn.domes((ny, nx), dx, [((nx - 1) * dx / 2, (ny - 1) * dx / 2)], [R], [300.0])
    if A is None:
        A = rng.uniform(50, 100, (ny, nx))
    if C is None:
        C = 0.0464
    return syn._pack(bed, H, A, C, dx)


def smooth_velocity(nz, ny, nx, seed=0, scale=100.0):
    """A smooth, level-dependent (U, V) pair (float64 numpy), non-trivial shear."""
    rng = np.random.default_rng(seed)
    base_u = syn.grf((ny, nx), 1.0, 4.0, 1.0, seed + 1) if min(ny, nx) > 2 else rng.standard_normal((ny, nx))
    base_v = syn.grf((ny, nx), 1.0, 4.0, 1.0, seed + 2) if min(ny, nx) > 2 else rng.standard_normal((ny, nx))
    z = np.linspace(0, 1, nz)[:, None, None]
    U = scale * (0.5 + base_u[None]) * (0.3 + z ** 2) + 5 * rng.standard_normal((nz, ny, nx))
    V = scale * (0.2 + base_v[None]) * (0.3 + z) + 5 * rng.standard_normal((nz, ny, nx))
    return U, V


def to_dev(a, dev, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=dev)


def rel(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
