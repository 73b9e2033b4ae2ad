"""Host helpers around the state fields: device placement and the fused
mask + derived-velocity kernel.

Reference: igm/processes/iceflow/utils.py:15-64 (initialize_iceflow_fields,
get_velbase/get_velsurf/get_velbar), solve/solve.py:116-117 and
emulate/emulate.py:143-152 (thk>0 mask).
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .energy import velocity_strides

OUT_NAMES = ("uvelbase", "vvelbase", "uvelsurf", "vvelsurf", "ubar", "vbar")


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("igm_amd needs a GPU: no HIP device is visible")
    return torch.device("cuda", torch.cuda.current_device())


def as_device(x, dtype=torch.float32):
    """numpy / torch (any device) -> contiguous device tensor of ``dtype``."""
    dev = device()
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x), dtype=dtype, device=dev)


def velocity_outputs(U, V, thk, vert_weight, nz, write_uv=True):
    """Masked planar U, V and the six derived (Ny, Nx) fields, one kernel.

    U, V: planar (Nz, Ny, Nx) pair, or Y (1, Ny, Nx, 2Nz) with V=None.
    Returns dict(U, V, uvelbase, vvelbase, uvelsurf, vvelsurf, ubar, vbar).
    """
    dt = thk.dtype
    Ny, Nx = thk.shape
    up, vp, _, ks, rs, cs = velocity_strides(U, V, nz)
    out = {n: torch.empty((Ny, Nx), dtype=dt, device=thk.device) for n in OUT_NAMES}
    Uo = torch.empty((nz, Ny, Nx), dtype=dt, device=thk.device) if write_uv else None
    Vo = torch.empty((nz, Ny, Nx), dtype=dt, device=thk.device) if write_uv else None
    vw = (C.c_double * nz)(*[float(w) for w in np.asarray(vert_weight).ravel()[:nz]])
    sfx = "f32" if dt == torch.float32 else "f64"
    L.call(f"igm_velocity_outputs_{sfx}", nz, Ny, Nx, up, vp, ks, rs, cs, L.ptr(thk),
           C.addressof(vw), L.ptr(Uo), L.ptr(Vo), *[L.ptr(out[n]) for n in OUT_NAMES],
           L.stream_ptr())
    out["U"], out["V"] = Uo, Vo
    return out
