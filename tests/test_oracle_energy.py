"""Pins of the CPU oracle's energy (oracle/energy.py), independent of its code:
closed forms of continuum mechanics, the textbook inclined slab, brute-force
finite differences, symmetry and convexity.  CPU only."""
import numpy as np
import pytest
import torch
from scipy.optimize import minimize

from oracle import energy as E
from oracle import vert_disc as VD

RHO_G = 910.0 * 9.81e-6  # MPa / m
NZ = 10


def dz64(nz=NZ):
    return torch.as_tensor(VD.compute_zeta_dzeta(VD.compute_levels(nz, 4.0))[1], dtype=torch.float64)


def fields(thk, usurf, A=78.0, C=0.0464, dx=100.0):
    ny, nx = np.shape(thk)
    return E.fieldin_tensors(thk, usurf, np.full((ny, nx), A), np.full((ny, nx), C), np.full((ny, nx), dx),
                             torch.float64)


def flat(ny, nx, H=300.0, slope_x=0.0, dx=100.0):
    x = np.arange(nx) * dx
    bed = 2500.0 - slope_x * np.broadcast_to(x[None, :], (ny, nx))
    thk = np.full((ny, nx), H)
    return thk, bed + thk


def test_zero_velocity_closed_form():
    """U=V=0: shear and gravity are exactly 0; sliding = C (regu_w^2)^(s/2) / s."""
    thk, us = flat(6, 7, slope_x=0.03)
    fin = fields(thk, us)
    z = torch.zeros((NZ, 6, 7), dtype=torch.float64)
    _, means = E.total_cost(z[None], z[None], fin, dz64())
    assert float(means[0]) == 0.0 and float(means[2]) == 0.0
    assert float(means[1]) == pytest.approx(0.0464 * (1e-20) ** (2 / 3) / (4 / 3), rel=1e-12)


def test_uniform_velocity_no_horizontal_strain():
    """A velocity uniform in x, y and z on a flat bed has zero strain:
    shear = B H (regu^2)^(-1/3) * 0 = 0 exactly."""
    thk, us = flat(5, 5)
    fin = fields(thk, us)
    U = torch.full((NZ, 5, 5), 42.0, dtype=torch.float64)
    maps = E.iceflow_energy(U[None], (0 * U)[None], fin, dz64())
    assert float(maps[0].abs().max()) == 0.0


@pytest.mark.parametrize("kind", ["extension", "shear"])
def test_effective_strain_rate_closed_form(kind):
    """Shear energy of a homogeneous strain field equals
    H * B * (3/4) * (eps_e^2 + regu^2)^(-1/3) * eps_e^2 with the effective
    strain rate eps_e^2 = 1/2 tr(eps^2) of the incompressible tensor:
    uniaxial extension u = e x: eps = diag(e, 0, -e) -> eps_e^2 = e^2;
    simple shear u = g y: eps_xy = g/2 -> eps_e^2 = g^2/4."""
    ny, nx, dx = 4, 5, 100.0
    thk, us = flat(ny, nx, H=250.0, dx=dx)
    fin = fields(thk, us, dx=dx)
    e = 0.003
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx)
    prof = X if kind == "extension" else Y
    U = torch.as_tensor(np.broadcast_to(e * prof, (NZ, ny, nx)).copy())
    eps2 = e * e if kind == "extension" else e * e / 4
    B = 2.0 * 78.0 ** (-1.0 / 3.0)
    expect = 250.0 * B * 0.75 * (eps2 + 1e-10) ** (-1.0 / 3.0) * eps2
    m = E.cost_shear(U[None], (0 * U)[None], *fin, dz64(), E.DEFAULT_PHYSICS)
    assert float(m.min()) == pytest.approx(expect, rel=1e-10)
    assert float(m.max()) == pytest.approx(expect, rel=1e-10)


def test_gravity_and_sliding_closed_forms():
    """Uniform u0 on a plane of surface slope -a: gravity = -rho g H a u0;
    flat-bed sliding = C (u0^2 + regu_w^2)^(2/3) / (4/3)."""
    ny, nx, dx = 4, 6, 100.0
    a, H, u0 = 0.04, 300.0, 37.0
    thk, us = flat(ny, nx, H=H, slope_x=a, dx=dx)
    fin = fields(thk, us, dx=dx)
    U = torch.full((NZ, ny, nx), u0, dtype=torch.float64)
    g = E.cost_gravity(U[None], (0 * U)[None], fin[0], fin[1], fin[4], dz64(), E.DEFAULT_PHYSICS)
    assert float(g.mean()) == pytest.approx(-RHO_G * H * a * u0, rel=1e-10)
    thk0, us0 = flat(ny, nx, H=H)
    fin0 = fields(thk0, us0, dx=dx)
    s = E.cost_sliding_weertman(U[None], (0 * U)[None], fin0[0], fin0[1], fin0[3], fin0[4], E.DEFAULT_PHYSICS)
    assert float(s.mean()) == pytest.approx(0.0464 * (u0 * u0 + 1e-20) ** (2 / 3) / (4 / 3), rel=1e-10)


@pytest.mark.parametrize("alpha,H,tol", [(0.02, 200.0, 0.008), (0.02, 400.0, 0.008), (0.05, 200.0, 0.025)])
def test_inclined_slab_minimiser(alpha, H, tol):
    """Textbook parallel-sided slab (Glen n=3, Weertman m=3):
    u(z) = (rho g H a / C)^3 + (A/2) (rho g a)^3 (H^4 - (H - z)^4).
    The discrete minimiser (one cell, per-level unknowns, L-BFGS in float64)
    matches it to O(a^2) (bed-following coordinates) + O(dzeta^2)."""
    nz, dx, A, C = NZ, 100.0, 78.0, 0.0464
    thk, us = flat(2, 2, H=H, slope_x=alpha, dx=dx)
    fin = fields(thk, us, A=A, C=C, dx=dx)
    dz = dz64()

    def f(u):
        U = torch.as_tensor(u).reshape(nz, 1, 1).expand(nz, 2, 2).clone().requires_grad_(True)
        c, _ = E.total_cost(U[None], torch.zeros_like(U)[None], fin, dz)
        (g,) = torch.autograd.grad(c, U)
        return float(c.detach()), g.sum(dim=(1, 2)).numpy()

    r = minimize(f, np.full(nz, 10.0), jac=True, method="L-BFGS-B",
                 options=dict(maxiter=5000, gtol=1e-16, ftol=1e-18))
    z = VD.compute_levels(nz, 4.0).astype(float) * H
    u = (RHO_G * alpha * H / C) ** 3 + (A / 2) * (RHO_G * alpha) ** 3 * (H ** 4 - (H - z) ** 4)
    assert np.max(np.abs(r.x - u) / u) <= tol


def _rand_problem(nz=5, ny=7, nx=9, seed=0):
    rng = np.random.default_rng(seed)
    dx = 100.0
    x = np.arange(nx) * dx
    bed = 2500 - 0.05 * x[None, :] + rng.normal(0, 5, (ny, nx))
    thk = np.maximum(0, 200 + rng.normal(0, 50, (ny, nx)))
    thk[0, 0] = 0.0
    fin = E.fieldin_tensors(thk, bed + thk, rng.uniform(50, 100, (ny, nx)), rng.uniform(0.01, 0.1, (ny, nx)),
                            np.full((ny, nx), dx), torch.float64)
    U = torch.as_tensor(rng.normal(50, 20, (nz, ny, nx)))
    V = torch.as_tensor(rng.normal(0, 20, (nz, ny, nx)))
    return fin, U, V, dz64(nz)


def test_gradient_central_differences():
    """Brute force: every 7th entry of dJ/dU, dJ/dV vs central differences."""
    fin, U, V, dz = _rand_problem()
    _, _, gU, gV = E.energy_and_grad(U, V, fin, dz)
    for T, G in ((U, gU), (V, gV)):
        flatT = T.view(-1)
        for idx in range(0, flatT.numel(), 7):
            h = 1e-4
            o = float(flatT[idx])
            flatT[idx] = o + h
            cp, _ = E.total_cost(U[None], V[None], fin, dz)
            flatT[idx] = o - h
            cm, _ = E.total_cost(U[None], V[None], fin, dz)
            flatT[idx] = o
            fd = (float(cp) - float(cm)) / (2 * h)
            assert fd == pytest.approx(float(G.view(-1)[idx]), rel=1e-6, abs=1e-12)


def test_transpose_symmetry():
    """Swapping x <-> y together with U <-> V leaves J unchanged."""
    fin, U, V, dz = _rand_problem(ny=8, nx=8, seed=3)
    c1, _ = E.total_cost(U[None], V[None], fin, dz)
    finT = [f.transpose(-1, -2).contiguous() for f in fin]
    c2, _ = E.total_cost(V.transpose(-1, -2)[None], U.transpose(-1, -2)[None], finT, dz)
    assert float(c1) == pytest.approx(float(c2), rel=1e-13)


def test_convexity_along_lines():
    """J(U + t D) has non-negative second differences (p = s = 4/3 > 1)."""
    fin, U, V, dz = _rand_problem(seed=5)
    rng = np.random.default_rng(9)
    for _ in range(5):
        DU = torch.as_tensor(rng.normal(0, 10, U.shape))
        DV = torch.as_tensor(rng.normal(0, 10, V.shape))
        ts = np.linspace(-2, 2, 21)
        Js = [float(E.total_cost((U + t * DU)[None], (V + t * DV)[None], fin, dz)[0]) for t in ts]
        d2 = np.diff(Js, 2)
        assert np.all(d2 >= -1e-12 * max(abs(j) for j in Js))


def test_float32_matches_float64():
    fin, U, V, dz = _rand_problem(seed=7)
    fin32 = [f.float() for f in fin]
    c64, _ = E.total_cost(U[None], V[None], fin, dz)
    c32, _ = E.total_cost(U.float()[None], V.float()[None], fin32, dz.float())
    assert float(c32) == pytest.approx(float(c64), rel=1e-5)


def _shear_slab(C, H=300.0, ax=0.0, ay=0.0, gamma=0.02, dx=100.0):
    """A slab of uniform thickness H on a plane bed of slopes (ax, ay), with
    the bed-parallel simple shear u = gamma (z - b): U_k = gamma H levels_k,
    V = 0.  Returns (fieldin, U, V)."""
    ny, nx = C.shape
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx)
    bed = 2500.0 + ax * X + ay * Y
    thk = np.full((ny, nx), H)
    fin = E.fieldin_tensors(thk, bed + thk, np.full((ny, nx), 78.0), C, np.full((ny, nx), dx), torch.float64)
    lev = torch.as_tensor(VD.compute_levels(NZ, 4.0), dtype=torch.float64)
    U = (gamma * H * lev)[:, None, None].expand(NZ, ny, nx).clone()
    return fin, U, torch.zeros_like(U)


@pytest.mark.parametrize("ax,ay", [(0.0, 0.0), (0.05, 0.0), (0.03, -0.04)])
def test_floating_damping_closed_form(ax, ay):
    """cost_shear.py:85-90: where the cell average of slidingco is <= 0, the
    vertical derivatives are scaled by 0.01 before the change of coordinate
    z -> z - b.  For the bed-parallel simple shear u = g (z - b) the
    Blatter-Pattyn strain-rate tensor has e_xz = g/2, e_xx = -e_zz = -g bx,
    e_xy = -g by / 2, so eps_e^2 = g^2 (bx^2 + by^2/4 + 1/4), and each cell's
    shear energy is H B (3/4) (eps_e^2 + regu^2)^(-1/3) eps_e^2 with g the
    shear rate, times 0.01 in the damped cells.  The sliding field has a
    block of zeros, a block of negative values and mixed-sign edges, so cells
    with some positive corners but a non-positive average are damped, and
    cells with a zero corner but a positive average are not."""
    ny, nx = 9, 11
    C = np.full((ny, nx), 0.0464)
    C[2:5, 3:7] = 0.0          # zero block: interior cells have average 0 -> damped
    C[6:9, 0:3] = -0.05        # negative block
    C[6, 3] = 0.01             # one positive corner next to the negative block
    C[0, 10] = -0.2            # one negative corner outweighs three positive ones: damped
    gamma, H = 0.02, 300.0
    fin, U, V = _shear_slab(C, H=H, ax=ax, ay=ay, gamma=gamma)
    m = E.cost_shear(U[None], V[None], *fin, dz64(), E.DEFAULT_PHYSICS)[0].numpy()
    B = 2.0 * 78.0 ** (-1.0 / 3.0)
    ndamp = nund = 0
    for j in range(ny - 1):
        for i in range(nx - 1):
            cs = (C[j, i] + C[j, i + 1] + C[j + 1, i] + C[j + 1, i + 1]) / 4
            g = gamma * (1.0 if cs > 0 else 0.01)
            eps2 = g * g * (ax * ax + ay * ay / 4 + 0.25)
            expect = H * B * 0.75 * (eps2 + 1e-10) ** (-1.0 / 3.0) * eps2
            assert m[j, i] == pytest.approx(expect, rel=1e-10), (j, i, cs)
            if cs > 0:
                nund += 1
            else:
                ndamp += 1
    assert ndamp >= 10 and nund >= 10


def test_floating_damping_scales_shear_like_slower_ice():
    """With no slope terms the damped cell's shear energy equals the undamped
    energy of the field scaled by 0.01, and with regu_glen small against the
    strain rate the ratio is the power law's 0.01^(4/3)."""
    ny, nx = 4, 4
    fin_g, U, V = _shear_slab(np.full((ny, nx), 0.0464), gamma=0.1)
    fin_f, _, _ = _shear_slab(np.zeros((ny, nx)), gamma=0.1)
    dz = dz64()
    grounded = E.cost_shear(U[None], V[None], *fin_g, dz, E.DEFAULT_PHYSICS)
    floating = E.cost_shear(U[None], V[None], *fin_f, dz, E.DEFAULT_PHYSICS)
    slow = E.cost_shear(0.01 * U[None], V[None], *fin_g, dz, E.DEFAULT_PHYSICS)
    assert torch.allclose(floating, slow, rtol=1e-13, atol=0)
    # (regu_glen^2 = 1e-10 against the damped eps_e^2 = 2.5e-7 moves the ratio by ~1.3e-4)
    assert float(floating.mean() / grounded.mean()) == pytest.approx(0.01 ** (4.0 / 3.0), rel=5e-4)
