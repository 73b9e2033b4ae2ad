"""Discretised Blatter-Pattyn energy (staggered grid, Lagrange basis).

torch-CPU restatement of ``igm/processes/iceflow/energy/`` in the reference's
op order, in float32 or float64 (the dtype of the inputs).  Autograd of these
functions stands in for the reference's ``tf.GradientTape``.

Shapes follow the reference: U, V are (1, Nz, Ny, Nx); thk, usurf, arrhenius,
slidingco, dX are (1, Ny, Nx); each cost returns a (1, Ny-1, Nx-1) map.

Pins (``tests/test_oracle_energy.py``): U=V=0 closed form, exact vanishing of
horizontal derivatives for uniform fields, thk=0 maps, the inclined-slab
minimiser, x<->y transpose symmetry, convexity along lines, and central finite
differences of the float64 energy against its autograd gradient.
"""
import math

import torch

DEFAULT_PHYSICS = dict(
    # conf/processes/iceflow.yaml:5-22
    exp_glen=3.0,
    exp_weertman=3.0,
    regu_glen=1.0e-05,
    regu_weertman=1.0e-10,
    thr_ice_thk=0.1,
    min_sr=1.0e-20,
    max_sr=1.0e20,
    ice_density=910.0,
    gravity_cst=9.81,
    force_negative_gravitational_energy=False,
)


def stag4h(B):
    """energy/utils.py:21-25 -- 4-point horizontal average onto cell centres."""
    return (B[..., 1:, 1:] + B[..., 1:, :-1] + B[..., :-1, 1:] + B[..., :-1, :-1]) / 4


def stag2v(B):
    """energy/utils.py:15-19 -- 2-point vertical average onto layer midpoints."""
    if B.shape[-3] > 1:
        B = (B[..., :-1, :, :] + B[..., 1:, :, :]) / 2
    return B


def compute_gradient_stag(s, dx, dy):
    """utils/gradient/compute_gradient.py:6-16 -- staggered slope of s."""
    E = 2.0 * (s[..., :, 1:] - s[..., :, :-1]) / (dx[..., :, 1:] + dx[..., :, :-1])
    diffx = 0.5 * (E[..., 1:, :] + E[..., :-1, :])
    EE = 2.0 * (s[..., 1:, :] - s[..., :-1, :]) / (dy[..., 1:, :] + dy[..., :-1, :])
    diffy = 0.5 * (EE[..., :, 1:] + EE[..., :, :-1])
    return diffx, diffy


def compute_horizontal_derivatives(U, V, dx):
    """energy/cost_shear.py:26-38 (staggered branch) -- level derivatives."""
    dUdx = (U[..., :, :, 1:] - U[..., :, :, :-1]) / dx
    dVdx = (V[..., :, :, 1:] - V[..., :, :, :-1]) / dx
    dUdy = (U[..., :, 1:, :] - U[..., :, :-1, :]) / dx
    dVdy = (V[..., :, 1:, :] - V[..., :, :-1, :]) / dx
    dUdx = (dUdx[..., :, :-1, :] + dUdx[..., :, 1:, :]) / 2
    dVdx = (dVdx[..., :, :-1, :] + dVdx[..., :, 1:, :]) / 2
    dUdy = (dUdy[..., :, :, :-1] + dUdy[..., :, :, 1:]) / 2
    dVdy = (dVdy[..., :, :, :-1] + dVdy[..., :, :, 1:]) / 2
    return dUdx, dVdx, dUdy, dVdy


def compute_srxy2(dUdx, dVdx, dUdy, dVdy):
    """energy/cost_shear.py:53-61."""
    Exx = dUdx
    Eyy = dVdy
    Ezz = -dUdx - dVdy
    Exy = 0.5 * dVdx + 0.5 * dUdy
    return 0.5 * (Exx**2 + Exy**2 + Exy**2 + Eyy**2 + Ezz**2)


def compute_srz2(dUdz, dVdz):
    """energy/cost_shear.py:63-69."""
    Exz = 0.5 * dUdz
    Eyz = 0.5 * dVdz
    return 0.5 * (Exz**2 + Eyz**2 + Exz**2 + Eyz**2)


def compute_vertical_derivatives(U, V, thk, dzeta, thr):
    """energy/cost_shear.py:71-83 (U, V, thk already on cell centres)."""
    if U.shape[-3] > 1:
        den = dzeta[None, :, None, None] * torch.clamp(thk, min=thr)[:, None, :, :]
        dUdz = (U[:, 1:, :, :] - U[:, :-1, :, :]) / den
        dVdz = (V[:, 1:, :, :] - V[:, :-1, :, :]) / den
    else:
        dUdz = torch.zeros_like(U)
        dVdz = torch.zeros_like(V)
    return dUdz, dVdz


def cost_shear(U, V, thk, usurf, arrhenius, slidingco, dX, dzeta, phys):
    """energy/cost_shear.py:103-179 (staggered_grid=1, vert_basis Lagrange).

    Note (cost_shear.py:172): the clip bounds are min_sr**2, max_sr**2 cast to
    the working dtype; in float32 max_sr**2 = 1e40 overflows to +inf, so only
    the lower clip is active, as in the reference.
    """
    n = phys["exp_glen"]
    B = 2.0 * arrhenius ** (-1.0 / n)
    if B.dim() == 3:
        B = B[:, None, :, :]
    p = 1.0 + 1.0 / n

    dx = dX[0, 0, 0]
    dUdx, dVdx, dUdy, dVdy = compute_horizontal_derivatives(U, V, dx)
    sloptopgx, sloptopgy = compute_gradient_stag(usurf - thk, dX, dX)

    U = stag4h(U)
    V = stag4h(V)
    slidingco = stag4h(slidingco)
    thk = stag4h(thk)
    B = stag4h(B)

    dUdx = stag2v(dUdx)
    dVdx = stag2v(dVdx)
    dUdy = stag2v(dUdy)
    dVdy = stag2v(dVdy)
    B = stag2v(B)

    dUdz, dVdz = compute_vertical_derivatives(U, V, thk, dzeta, phys["thr_ice_thk"])
    # cost_shear.py:85-90: dampen where the (staggered) sliding coefficient <= 0
    floating = (slidingco[:, None, :, :] > 0)
    dUdz = torch.where(floating, dUdz, 0.01 * dUdz)
    dVdz = torch.where(floating, dVdz, 0.01 * dVdz)
    # cost_shear.py:92-101: change of coordinate z -> z - b
    dUdx = dUdx - dUdz * sloptopgx[:, None, :, :]
    dUdy = dUdy - dUdz * sloptopgy[:, None, :, :]
    dVdx = dVdx - dVdz * sloptopgx[:, None, :, :]
    dVdy = dVdy - dVdz * sloptopgy[:, None, :, :]

    sr2 = compute_srxy2(dUdx, dVdx, dUdy, dVdy) + compute_srz2(dUdz, dVdz)
    lo = torch.tensor(phys["min_sr"] ** 2, dtype=sr2.dtype)
    hi = torch.tensor(phys["max_sr"] ** 2, dtype=sr2.dtype)
    sr2capped = torch.clamp(sr2, min=lo, max=hi)
    p_term = ((sr2capped + phys["regu_glen"] ** 2) ** ((p - 2) / 2)) * sr2 / p
    return thk * torch.sum(B * dzeta[None, :, None, None] * p_term, dim=1)


def cost_gravity(U, V, thk, usurf, dX, dzeta, phys):
    """energy/cost_gravity.py:24-69 (staggered_grid=1, vert_basis Lagrange)."""
    slopsurfx, slopsurfy = compute_gradient_stag(usurf, dX, dX)
    U = stag4h(U)
    V = stag4h(V)
    thk = stag4h(thk)
    U = stag2v(U)
    V = stag2v(V)
    uds = U * slopsurfx[:, None, :, :] + V * slopsurfy[:, None, :, :]
    if phys["force_negative_gravitational_energy"]:
        uds = torch.clamp(uds, max=0.0)
    return (phys["ice_density"] * phys["gravity_cst"] * 10 ** (-6) * thk
            * torch.sum(dzeta[None, :, None, None] * uds, dim=1))


def cost_sliding_weertman(U, V, thk, usurf, slidingco, dX, phys):
    """energy/cost_sliding_weertman.py:23-43 (staggered_grid=1, Lagrange)."""
    C = 1.0 * slidingco
    s = 1.0 + 1.0 / phys["exp_weertman"]
    sloptopgx, sloptopgy = compute_gradient_stag(usurf - thk, dX, dX)
    U = stag4h(U)
    V = stag4h(V)
    C = stag4h(C)
    uvelbase, vvelbase = U[..., 0, :, :], V[..., 0, :, :]   # iceflow/utils.py:34-36
    N = ((uvelbase ** 2 + vvelbase ** 2) + phys["regu_weertman"] ** 2
         + (uvelbase * sloptopgx + vvelbase * sloptopgy) ** 2)
    return C * N ** (s / 2) / s


def iceflow_energy(U, V, fieldin, dzeta, phys=None):
    """energy/energy.py:12-22 with the default energy_components
    ['shear', 'sliding_weertman', 'gravity'] (conf/processes/iceflow.yaml:5).

    ``fieldin`` = [thk, usurf, arrhenius, slidingco, dX], each (1, Ny, Nx).
    Returns the list of per-cell maps in that component order.
    """
    phys = DEFAULT_PHYSICS if phys is None else phys
    thk, usurf, arrhenius, slidingco, dX = fieldin
    dzeta = torch.as_tensor(dzeta, dtype=U.dtype)
    return [
        cost_shear(U, V, thk, usurf, arrhenius, slidingco, dX, dzeta, phys),
        cost_sliding_weertman(U, V, thk, usurf, slidingco, dX, phys),
        cost_gravity(U, V, thk, usurf, dX, dzeta, phys),
    ]


def total_cost(U, V, fieldin, dzeta, phys=None):
    """solve/solve.py:50-59: COST = add_n([reduce_mean(e) for e in maps])."""
    maps = iceflow_energy(U, V, fieldin, dzeta, phys)
    means = [torch.mean(m) for m in maps]
    return means[0] + means[1] + means[2], means


def energy_and_grad(U, V, fieldin, dzeta, phys=None):
    """Energy and its gradient w.r.t. (U, V) -- the tape of solve.py:45-77.

    U, V: (Nz, Ny, Nx) tensors (no batch dim), any float dtype.
    Returns (COST, [shear, sliding, gravity] means, dU, dV).
    """
    Uv = U.detach().clone().requires_grad_(True)
    Vv = V.detach().clone().requires_grad_(True)
    cost, means = total_cost(Uv[None], Vv[None], fieldin, dzeta, phys)
    gU, gV = torch.autograd.grad(cost, [Uv, Vv])
    return cost.detach(), [m.detach() for m in means], gU, gV


def fieldin_tensors(thk, usurf, arrhenius, slidingco, dX, dtype=torch.float32):
    """Build the (1, Ny, Nx) fieldin list the energy expects (solve.py:32)."""
    return [torch.as_tensor(f, dtype=dtype)[None] for f in (thk, usurf, arrhenius, slidingco, dX)]


def rho_g(phys=None):
    phys = DEFAULT_PHYSICS if phys is None else phys
    return phys["ice_density"] * phys["gravity_cst"] * 10 ** (-6)


def zero_field_cost(slidingco_mean_stag, phys=None):
    """Closed form of J at U=V=0: only the sliding term survives,
    mean(C_s * (regu_weertman**2)**(s/2) / s)."""
    phys = DEFAULT_PHYSICS if phys is None else phys
    s = 1.0 + 1.0 / phys["exp_weertman"]
    return slidingco_mean_stag * math.pow(phys["regu_weertman"] ** 2, s / 2) / s
