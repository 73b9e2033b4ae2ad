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


def misfits(thk, U_ref, V_ref, U, V, vert_weight, vert_spacing):
    """The reference's accuracy measures of a velocity field (U, V) against
    (U_ref, V_ref), all planar (Nz, Ny, Nx) on any device, in float64:

    * ``rel_l1``, ``rel_l2``: pretraining.py:502-535 (_computemisfitall) --
      the thickness-weighted L1 and L2 norms of the layer-midpoint velocity
      difference over the ice column, relative to those of the reference
      field; the layer thicknesses are thk times the steps of the stretched
      vertical coordinate (vert_disc.py:10-13);
    * ``velbar_l1``, ``velbar_rms``: diagnostic.py:74-84 (computemisfit) --
      the mean absolute and root-mean-square difference of the
      depth-averaged velocity over cells with thk > 1.

    A measurement for reports (float64 torch reductions), not a hot path.
    """
    nz = U.shape[0]
    f64 = torch.float64
    thk, U_ref, V_ref, U, V = (t.to(f64) for t in (thk, U_ref, V_ref, U, V))
    zeta = torch.arange(nz, dtype=f64, device=U.device) / (nz - 1)
    lev = (zeta / vert_spacing) * (1.0 + (vert_spacing - 1.0) * zeta)
    dz = (lev[1:] - lev[:-1])[:, None, None] * thk[None]

    def mid(a):
        return 0.5 * (a[1:] + a[:-1])

    du, dv = mid(U_ref) - mid(U), mid(V_ref) - mid(V)
    ur, vr = mid(U_ref), mid(V_ref)

    def normp(a, b, p):
        return float(torch.sum(dz * (a.abs() ** p + b.abs() ** p)) ** (1.0 / p))

    w = torch.as_tensor(vert_weight).to(device=U.device, dtype=f64).reshape(-1)[:, None, None]
    dub, dvb = torch.sum(w * (U_ref - U), 0), torch.sum(w * (V_ref - V), 0)
    m = (thk > 1).to(f64)
    n = 2.0 * float(torch.sum(m))
    l1 = (float(torch.sum(dub.abs() * m) + torch.sum(dvb.abs() * m)) / n) if n else 0.0
    l2 = (float(torch.sum(dub ** 2 * m) + torch.sum(dvb ** 2 * m)) / n) if n else 0.0
    return {"rel_l1": normp(du, dv, 1.0) / normp(ur, vr, 1.0),
            "rel_l2": normp(du, dv, 2.0) / normp(ur, vr, 2.0),
            "velbar_l1": l1, "velbar_rms": float(np.sqrt(l2))}
