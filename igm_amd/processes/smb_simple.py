"""Drop-in ``smb_simple`` process (igm/processes/smb_simple/smb_simple.py:10-50):
the ELA-gradient surface mass balance, parameters interpolated in time on the
host (interp1d, utils/math/interp1d_tf.py:4-41), field evaluated by
igm_smb_simple_f32."""
import numpy as np
import torch

from .. import _lib as L


def initialize(cfg, state):
    """smb_simple.py:10-21: table rows (time, gradabl, gradacc, ela, maxacc)."""
    c = cfg.processes.smb_simple
    if len(c.array) == 0:
        state.smbpar = np.loadtxt(c.file, skiprows=1, dtype=np.float32)
    else:
        state.smbpar = np.array(c.array[1:]).astype(np.float32)
    state.smbpar = np.atleast_2d(state.smbpar)
    state.tlast_mb = -1.0e5000


def _interp(x, y, t):
    return float(np.interp(t, x, y))  # linear, clamped at the ends like interp1d_tf


def update(cfg, state):
    """smb_simple.py:24-50."""
    c = cfg.processes.smb_simple
    if (state.t - state.tlast_mb) >= c.update_freq:
        p = state.smbpar
        gradabl, gradacc, ela, maxacc = (_interp(p[:, 0], p[:, k], state.t) for k in (1, 2, 3, 4))
        if not hasattr(state, "smb") or state.smb.shape != state.usurf.shape:
            state.smb = torch.empty_like(state.usurf)
        L.call("igm_smb_simple_f32", state.usurf.numel(), L.ptr(state.usurf.contiguous()), gradabl, gradacc, ela,
               maxacc, L.ptr(state.smb), L.stream_ptr())
        if hasattr(state, "icemask"):
            raise NotImplementedError("icemask is not supported")
        state.tlast_mb = state.t


def finalize(cfg, state):
    pass
