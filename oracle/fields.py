"""Derived velocity fields, layout transforms and misfits (CPU restatement).

Restates ``igm/processes/iceflow/utils.py:34-64`` (get_vel*),
``utils.py:133-158`` (Y_to_UV / UV_to_Y / fieldin_to_X),
``solve/solve.py:116-117`` and ``emulate/emulate.py:143-152`` (thk>0 mask),
``diagnostic/diagnostic.py:74-84`` (computemisfit) and
``pretraining/pretraining.py:502-535`` (_computemisfitall).
Pinned by construction tests (weights summing to one, masks, round trips).
"""
import numpy as np
import torch


def mask_thk(U, thk):
    """solve.py:116-117 / emulate.py:143-144: where(thk > 0, U, 0)."""
    m = torch.as_tensor(thk) > 0
    return torch.where(m, U, torch.zeros_like(U))


def get_velbase(U, V):
    """utils.py:34-41 (Lagrange)."""
    return U[..., 0, :, :], V[..., 0, :, :]


def get_velsurf(U, V):
    """utils.py:43-51 (Lagrange)."""
    return U[..., -1, :, :], V[..., -1, :, :]


def get_velbar(U, V, vert_weight):
    """utils.py:53-64 (Lagrange): sum_k vert_weight_k U_k."""
    w = torch.as_tensor(vert_weight, dtype=U.dtype)
    return torch.sum(U * w, dim=-3), torch.sum(V * w, dim=-3)


def velocity_outputs(U, V, thk, vert_weight):
    """solve.py:116-117 + 186-188: mask then the six derived 2-D fields."""
    U = mask_thk(U, thk)
    V = mask_thk(V, thk)
    ub, vb = get_velbase(U, V)
    us, vs = get_velsurf(U, V)
    ubar, vbar = get_velbar(U, V, vert_weight)
    return dict(U=U, V=V, uvelbase=ub, vvelbase=vb, uvelsurf=us, vvelsurf=vs,
                ubar=ubar, vbar=vbar)


def fieldin_to_X(fieldin):
    """utils.py:147-158 (dim_arrhenius=2): stack to NHWC (1, Ny, Nx, 5)."""
    return torch.stack([torch.as_tensor(f) for f in fieldin], dim=-1)[None]


def Y_to_UV(Y, Nz):
    """utils.py:133-139: Y (1, Ny, Nx, 2Nz) -> U, V (1, Nz, Ny, Nx)."""
    U = torch.movedim(Y[..., :Nz], -1, 1)
    V = torch.movedim(Y[..., Nz:], -1, 1)
    return U, V


def UV_to_Y(U, V):
    """utils.py:141-145: U, V (Nz, Ny, Nx) -> Y (1, Ny, Nx, 2Nz)."""
    return torch.cat([torch.movedim(U, 0, -1), torch.movedim(V, 0, -1)], dim=-1)[None]


def computemisfit(thk, dU, dV, vert_weight):
    """diagnostic.py:74-84: masked (thk>1) mean L1 and RMS of the velbar misfit."""
    ubar, vbar = get_velbar(dU, dV, vert_weight)
    VEL = torch.stack([ubar, vbar], dim=0)
    MA = torch.where(torch.as_tensor(thk) > 1, torch.ones_like(VEL), torch.zeros_like(VEL))
    nl1 = torch.sum(MA * torch.abs(VEL)) / torch.sum(MA)
    nl2 = torch.sum(MA * torch.abs(VEL) ** 2) / torch.sum(MA)
    return float(nl1), float(np.sqrt(float(nl2)))


def rel_l2(a, b, mask=None):
    """Relative L2 error of a against b over the masked (ice) nodes.

    Not a reference function: the error metric used to compare velocity
    fields (a, b of shape (Nz, Ny, Nx) or stacked U/V)."""
    a = torch.as_tensor(a, dtype=torch.float64)
    b = torch.as_tensor(b, dtype=torch.float64)
    if mask is not None:
        m = torch.as_tensor(mask).to(torch.bool)
        a = a * m
        b = b * m
    return float(torch.linalg.vector_norm(a - b) / torch.clamp(torch.linalg.vector_norm(b), min=1e-300))


def computemisfitall(thk, Ut, Vt, Up, Vp, vert_weight, vert_spacing):
    """pretraining.py:502-535: (rel L1, rel L2) of the layer-midpoint velocity
    difference weighted by the layer thickness thk*(lev_{l+1}-lev_l), with the
    stretched levels of vert_disc.py:10-13, and computemisfit's (L1, RMS) of
    the velbar difference.  (The reference returns the velbar L1 twice in
    place of (L1, L2); the restatement returns both.)  Test infrastructure."""
    Nz = Ut.shape[0]
    thk = torch.as_tensor(thk, dtype=torch.float64)
    zeta = np.arange(Nz) / (Nz - 1)
    temp = (zeta / vert_spacing) * (1.0 + (vert_spacing - 1.0) * zeta)
    temp = temp[1:] - temp[:-1]
    dz = torch.stack([thk * float(z) for z in temp])
    Ut, Vt, Up, Vp = (torch.as_tensor(a, dtype=torch.float64) for a in (Ut, Vt, Up, Vp))
    l1bar, l2bar = computemisfit(thk, Ut - Up, Vt - Vp, vert_weight)

    def half(a):
        return 0.5 * (a[1:, :, :] + a[:-1, :, :])

    ut, up, vt, vp = half(Ut), half(Up), half(Vt), half(Vp)

    def normp(u, v, p):
        temp = torch.sum(dz * (torch.abs(u) ** p + torch.abs(v) ** p), dim=0)
        return float(torch.sum(temp) ** (1.0 / p))

    return (normp(ut - up, vt - vp, 1.0) / normp(ut, vt, 1.0),
            normp(ut - up, vt - vp, 2.0) / normp(ut, vt, 2.0), l1bar, l2bar)
