"""Drop-in ``thk`` process (igm/processes/thk/thk.py:11-58): the fused
slope-limited upwind thickness update on the GPU (igm_transport_f32)."""
import torch

from .. import _lib as L
from ..fields import as_device

SLOPES = {"godunov": 0, "minmod": 1, "superbee": 2}


def _sealevel(cfg, state):
    return float(getattr(state, "sealevel", cfg.processes.thk.default_sealevel))


def initialize(cfg, state):
    """thk.py:11-24: lsurf, usurf from topg and thk."""
    if not hasattr(state, "topg"):
        raise ValueError("The 'thk' module requires an initial topography ('state.topg') to be defined. "
                         "Please define it through the preprocessing steps (not yet implemented)")
    state.topg = as_device(state.topg)
    state.thk = as_device(state.thk)
    r = cfg.processes.thk.ratio_density
    state.lsurf = torch.maximum(state.topg, -r * state.thk + _sealevel(cfg, state))
    state.usurf = state.lsurf + state.thk


def update(cfg, state):
    """thk.py:27-54."""
    if state.it < 0:
        return
    if hasattr(state, "logger"):
        state.logger.info("Ice thickness equation at time : " + str(float(state.t)))
    c = cfg.processes.thk
    ny, nx = state.thk.shape
    if not hasattr(state, "smb"):
        state.smb = torch.zeros_like(state.thk)
    thk_new = torch.empty_like(state.thk)
    lsurf = torch.empty_like(state.thk)
    usurf = torch.empty_like(state.thk)
    divflux = torch.empty_like(state.thk)
    dx = float(state.dx)
    L.call("igm_transport_f32", ny, nx, L.ptr(state.ubar.contiguous()), L.ptr(state.vbar.contiguous()),
           L.ptr(state.thk), L.ptr(state.smb), L.ptr(state.topg), dx, dx, float(state.dt), SLOPES[c.slope_type],
           float(c.ratio_density), _sealevel(cfg, state), L.ptr(thk_new), L.ptr(lsurf), L.ptr(usurf),
           L.ptr(divflux), L.stream_ptr())
    state.thk, state.lsurf, state.usurf, state.divflux = thk_new, lsurf, usurf, divflux


def finalize(cfg, state):
    pass
