"""Pins of the committed dome100 minimiser (tests/golden/dome100_minimiser.npz,
written by tests/golden/make_dome100_minimiser.py from the oracle's float64
energy): the stored energy is the oracle's energy of the stored field; the
field is a critical point (gradient norm tiny against the one at zero); and
no step along random directions, nor along the negative gradient, lowers the
energy (the energy is convex, so a point that no small step improves is the
minimum).  CPU only."""
import os

import numpy as np
import pytest
import torch

from igm_amd import synthetic as syn
from oracle import energy as E
from oracle import vert_disc as VD

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dome100_minimiser.npz")


@pytest.fixture(scope="module")
def problem():
    d = np.load(FIX)
    g = syn.dome100()
    fin = E.fieldin_tensors(*(g[k] for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")), dtype=torch.float64)
    dz = torch.as_tensor(VD.compute_zeta_dzeta(VD.compute_levels(10, 4.0))[1], dtype=torch.float64)
    return d, fin, dz, g


def test_stored_energy_and_critical_point(problem):
    d, fin, dz, g = problem
    U, V = torch.as_tensor(d["U"]), torch.as_tensor(d["V"])
    c, _, gU, gV = E.energy_and_grad(U, V, fin, dz)
    assert float(c) == pytest.approx(float(d["J"]), rel=1e-13)
    z = torch.zeros_like(U)
    _, _, g0U, g0V = E.energy_and_grad(z, z, fin, dz)
    # on the ice nodes (the ice-free margin nodes carry the sliding term,
    # whose gradient falls only as |u|^(1/3) near their zero optimum and
    # dominates the norm elsewhere)
    ice = torch.as_tensor(g["thk"] > 0)
    G, G0 = torch.stack([gU, gV])[:, :, ice], torch.stack([g0U, g0V])[:, :, ice]
    ratio = float(G.norm() / G0.norm())
    assert ratio <= 1e-4, ratio


def test_no_small_step_lowers_the_energy(problem):
    d, fin, dz, g = problem
    U, V = torch.as_tensor(d["U"]), torch.as_tensor(d["V"])
    J0, _ = E.total_cost(U[None], V[None], fin, dz)
    J0 = float(J0)
    _, _, gU, gV = E.energy_and_grad(U, V, fin, dz)
    ice = torch.as_tensor(g["thk"] > 0)
    scale = float(torch.stack([U, V])[:, :, ice].abs().mean())
    rng = np.random.default_rng(0)
    dirs = [(-gU / gU.norm(), -gV / gU.norm())]
    for _ in range(4):
        DU = torch.as_tensor(rng.standard_normal(U.shape)) * ice
        DV = torch.as_tensor(rng.standard_normal(V.shape)) * ice
        n = float(torch.stack([DU, DV]).norm())
        dirs.append((DU / n, DV / n))
    for DU, DV in dirs:
        for h in (1e-2, 1e-1, 1.0):
            s = h * scale
            for sign in (1.0, -1.0):
                J, _ = E.total_cost((U + sign * s * DU)[None], (V + sign * s * DV)[None], fin, dz)
                # float64 sum over 10,000 cells: a rounding floor of ~1e-14 |J|
                assert float(J) >= J0 - 1e-13 * abs(J0), (h, sign, float(J) - J0)
