"""Drop-in ``vert_flow`` process (igm/processes/vert_flow/vert_flow.py:18-59):
the vertical velocity W (Nz, Ny, Nx) from state.U, state.V, topg and thk,
plus wvelbase / wvelsurf, computed by igm_vert_flow_kinematic_f32.

Supported: the default, Lagrange basis with version 2 / method "kinematic"
(vert_flow_v2.py:17-37).  The other variants (version 1, "incompressibility",
Legendre, SIA) raise NotImplementedError."""
import ctypes as C

import numpy as np
import torch

from .. import _lib as L
from ..vert_disc import compute_levels, define_vertical_weight

DEFAULTS = {"version": 2, "method": "kinematic"}


def initialize(cfg, state):
    """vert_flow.py:18-19 (nothing to set up)."""


def _check(cfg):
    vf = getattr(cfg.processes, "vert_flow", None)
    version = int(getattr(vf, "version", DEFAULTS["version"])) if vf is not None else DEFAULTS["version"]
    method = getattr(vf, "method", DEFAULTS["method"]) if vf is not None else DEFAULTS["method"]
    basis = cfg.processes.iceflow.numerics.vert_basis
    if basis not in ("Lagrange", "Legendre", "SIA"):
        raise ValueError(f"Unknown vertical basis: {basis}")  # vert_flow.py:52-53
    if basis != "Lagrange" or version != 2 or method != "kinematic":
        raise NotImplementedError("vert_flow: only the Lagrange basis with version 2, method 'kinematic' "
                                  "is supported")


def update(cfg, state):
    """vert_flow.py:21-56."""
    _check(cfg)
    ice = cfg.processes.iceflow
    nz, vs = int(ice.numerics.Nz), float(ice.numerics.vert_spacing)
    U, V = state.U.contiguous(), state.V.contiguous()
    assert U.dtype == torch.float32 and U.shape[0] == nz
    ny, nx = U.shape[1:]
    levels = (C.c_double * nz)(*[float(x) for x in np.asarray(compute_levels(nz, vs)).ravel()])
    vw = (C.c_double * nz)(*[float(x) for x in np.asarray(define_vertical_weight(nz, vs)).ravel()])
    W = torch.empty_like(U)
    wb = torch.empty((ny, nx), dtype=torch.float32, device=U.device)
    ws = torch.empty_like(wb)
    topg = state.topg.contiguous()
    thk = state.thk.contiguous()
    L.call("igm_vert_flow_kinematic_f32", nz, ny, nx, L.ptr(U), L.ptr(V), L.ptr(topg), L.ptr(thk),
           C.addressof(levels), C.addressof(vw), float(state.dx), L.ptr(W), L.ptr(wb), L.ptr(ws),
           L.stream_ptr())
    state.W, state.wvelbase, state.wvelsurf = W, wb, ws


def finalize(cfg, state):
    """vert_flow.py:58-59."""
