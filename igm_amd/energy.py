"""Energy operator: the fused HIP energy/adjoint pass behind a Python API.

``EnergyOperator`` holds the velocity-independent per-cell coefficients of
one problem (computed once by ``igm_energy_coeffs_*``) and evaluates the
Blatter-Pattyn energy of (U, V) -- and its gradient -- with
``igm_energy_pass_*`` + ``igm_energy_finalize``.  ``EnergyFunction`` wraps it
as a ``torch.autograd.Function`` whose forward pass already computes the
gradient (one fused sweep), so backward is a scale.

Reference: igm/processes/iceflow/energy/energy.py:12-30 (iceflow_energy,
iceflow_energy_XY), cost_shear.py, cost_gravity.py,
cost_sliding_weertman.py; solve/solve.py:50-77.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib as L

DTYPES = {torch.float32: "f32", torch.float64: "f64"}


def physics_struct(phys):
    """IgmPhysics from a cfg.processes.iceflow.physics-like object or dict."""
    g = (lambda k: phys[k]) if isinstance(phys, dict) else (lambda k: getattr(phys, k))
    if g("force_negative_gravitational_energy"):
        raise NotImplementedError("force_negative_gravitational_energy is not supported")
    return L.IgmPhysics(
        exp_glen=float(g("exp_glen")), exp_weertman=float(g("exp_weertman")),
        regu_glen=float(g("regu_glen")), regu_weertman=float(g("regu_weertman")),
        thr_ice_thk=float(g("thr_ice_thk")), min_sr=float(g("min_sr")), max_sr=float(g("max_sr")),
        ice_density=float(g("ice_density")), gravity_cst=float(g("gravity_cst")),
        force_negative_gravitational_energy=0)


def velocity_strides(U, V=None, nz=None):
    """(U ptr, V ptr, bstride, kstride, rstride, cstride) of a velocity pair.

    Accepts planar U, V of shape (Nz, Ny, Nx) or (B, Nz, Ny, Nx), or an
    emulator output Y of shape (B, Ny, Nx, 2Nz) (V=None, nz given).
    """
    if V is None:
        Y = U
        assert Y.dim() == 4 and nz is not None and Y.shape[-1] == 2 * nz
        es = Y.element_size()
        return (Y.data_ptr(), Y.data_ptr() + nz * es * Y.stride(3), Y.stride(0), Y.stride(3),
                Y.stride(1), Y.stride(2))
    assert U.shape == V.shape and U.stride() == V.stride()
    if U.dim() == 3:
        return (U.data_ptr(), V.data_ptr(), 0, U.stride(0), U.stride(1), U.stride(2))
    assert U.dim() == 4
    return (U.data_ptr(), V.data_ptr(), U.stride(0), U.stride(1), U.stride(2), U.stride(3))


def _check_pair_layout(op, U, V, Uo, Vo):
    """Shapes of a two-iteration pass's velocity buffers: planar (Nz, Ny, Nx)
    pairs, or with V (Vo) None an interleaved (Nz, Ny, Nx, 2) U (Uo)."""
    assert op.pair_supported
    for u, v in ((U, V), (Uo, Vo)):
        shape = (op.nz, op.Ny, op.Nx) + (() if v is not None else (2,))
        assert tuple(u.shape) == shape and u.is_contiguous(), (tuple(u.shape), shape)
        assert v is None or (tuple(v.shape) == shape and v.is_contiguous())


class EnergyOperator:
    """Coefficients + evaluation of the energy for a fixed set of input fields.

    fields: dict with thk, usurf, arrhenius, slidingco, dX as device tensors
    of shape (Ny, Nx) or (B, Ny, Nx); or ``X`` NHWC (B, Ny, Nx, 5) in the
    order [thk, usurf, arrhenius, slidingco, dX] (iceflow/utils.py:147-158).
    extents: optional int32 device tensor (B, 2) of valid (ny, nx) per entry.
    """

    def __init__(self, dzeta, phys, fields=None, X=None, extents=None, dtype=torch.float32):
        self.dtype = dtype
        self.sfx = DTYPES[dtype]
        self.nz = len(dzeta) + 1
        self.dzeta = (C.c_double * max(1, len(dzeta)))(*[float(d) for d in dzeta])
        self.phys = physics_struct(phys)
        if X is not None:
            assert X.dim() == 4 and X.shape[-1] == 5 and X.dtype == dtype and X.is_contiguous()
            B, Ny, Nx, _ = X.shape
            es = X.element_size()
            base = X.data_ptr()
            ptrs = [base + f * es for f in range(5)]
            fb, fr, fc = X.stride(0), X.stride(1), X.stride(2)
            self._keep = X
        else:
            fl = [fields[k] for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")]
            t0 = fl[0]
            if t0.dim() == 2:
                fl = [f[None] for f in fl]
            fl = [f.to(dtype).contiguous() for f in fl]
            B, Ny, Nx = fl[0].shape
            ptrs = [f.data_ptr() for f in fl]
            fb, fr, fc = fl[0].stride()
            self._keep = fl
        self.B, self.Ny, self.Nx = B, Ny, Nx
        dev = self._keep[0].device if isinstance(self._keep, list) else self._keep.device
        self.device = dev
        self.extents = extents
        self.coef = torch.empty((B, 7, Ny, Nx), dtype=dtype, device=dev)
        self.hscale = torch.empty((B,), dtype=dtype, device=dev)
        nws = L.load().igm_energy_workspace_doubles(B, Ny, Nx)
        self.partials = torch.zeros((max(1, nws),), dtype=torch.float64, device=dev)
        self._partials_b = None
        self.out = torch.zeros((B, 4), dtype=torch.float64, device=dev)
        L.call(f"igm_energy_coeffs_{self.sfx}", *ptrs, fb, fr, fc, L.ptr(extents), B, Ny, Nx,
               C.addressof(self.phys), L.ptr(self.coef), L.ptr(self.hscale), L.stream_ptr())

    @property
    def partials_b(self):
        """A second partials buffer (a pass's fused finalize reads the previous
        pass's partials while it writes its own)."""
        if self._partials_b is None:
            self._partials_b = torch.zeros_like(self.partials)
        return self._partials_b

    # -- raw passes ---------------------------------------------------------
    def run_pass(self, mode, U, V=None, gU=None, gV=None, Uo=None, Vo=None, moments=None, adam=None,
                 stop_flag=None, partials=None, prev=None):
        """One pass (mode 0 energy, 1 + gradient, 2 + Adam step).  ``moments``
        (mode 2): (B, Nz, Ny, Nx, 4) Adam slots (m_U, v_U, m_V, v_V).
        ``partials``: the per-block sums buffer (default: the operator's own).
        ``prev`` (optional): (prev_partials, es_in, es_out, cost_hist,
        prev_iter) -- the previous iteration's finalize, done by the pass
        itself (see ``solver.solve``)."""
        up, vp, bs, ks, rs, cs = velocity_strides(U, V, self.nz)
        if gU is not None and gV is None:  # gradient in Y layout
            gu, gv = velocity_strides(gU, None, self.nz)[:2]
        else:
            gu, gv = L.ptr(gU), L.ptr(gV)
        L.call(f"igm_energy_pass_{self.sfx}", mode, self.nz, self.B, self.Ny, self.Nx, up, vp,
               bs, ks, rs, cs, L.ptr(self.coef), L.ptr(self.hscale), C.addressof(self.dzeta),
               C.addressof(self.phys), gu, gv, L.ptr(Uo), L.ptr(Vo), L.ptr(moments),
               C.addressof(adam) if adam is not None else None,
               L.ptr(self.partials if partials is None else partials), stop_flag,
               *((L.ptr(prev[0]), prev[1], prev[2], L.ptr(self.out), L.ptr(prev[3]), prev[4])
                 if prev is not None else (None, None, None, None, None, 0)),
               L.stream_ptr())

    # Nz with a two-iteration pass (igm_energy_pass2_f32)
    PAIR_NZ = (2, 3, 4, 5, 6, 8, 10, 12)

    @property
    def pair_supported(self):
        """Whether ``run_pass2`` serves this operator: float32, one field and
        an Nz the two-iteration pass is compiled for."""
        return self.dtype == torch.float32 and self.B == 1 and self.nz in self.PAIR_NZ

    def run_pass2(self, U, V, Uo, Vo, mom_in, mom_out, adam1, adam2, partials, stop_flag=None,
                  prev=None):
        """Two solver iterations in one pass (planar (Nz, Ny, Nx) float32):
        energy, gradient and legacy-Adam step with ``adam1``, then again with
        ``adam2``.  Moments (Nz, Ny, Nx, 4) are read from ``mom_in`` and
        written to ``mom_out``.  ``partials`` receives 6 doubles per block.
        ``prev``: (prev_partials, prev_count, es_in, es_out, cost_hist,
        prev_iter) -- the previous pass's finalize, done by this pass."""
        _check_pair_layout(self, U, V, Uo, Vo)
        L.call("igm_energy_pass2_f32", self.nz, self.Ny, self.Nx, L.ptr(U), L.ptr(V),
               L.ptr(self.coef), L.ptr(self.hscale), C.addressof(self.dzeta), C.addressof(self.phys),
               L.ptr(Uo), L.ptr(Vo), L.ptr(mom_in), L.ptr(mom_out), C.addressof(adam1),
               C.addressof(adam2), L.ptr(partials), stop_flag,
               *((L.ptr(prev[0]), prev[1], prev[2], prev[3], L.ptr(self.out), L.ptr(prev[4]), prev[5])
                 if prev is not None else (None, 0, None, None, None, None, 0)),
               L.stream_ptr())

    def run_solve_pass(self, U, V, Uo, Vo, mom_in, mom_out, adam, alpha_tab, nbitmax, partials,
                       prev_partials, es_in, es_out, cost_hist, finalize_only=False):
        """One device-driven solve pass (``igm_energy_solve_pass_f32``): it
        accounts the previous pass and runs 0, 1 or 2 iterations as the
        EarlyStopping state allows.  ``alpha_tab``: float32 device tensor of
        the per-iteration Adam steps; ``es_in``, ``es_out``: device addresses
        of two IgmEarlyStop states."""
        _check_pair_layout(self, U, V, Uo, Vo)
        L.call("igm_energy_solve_pass_f32", self.nz, self.Ny, self.Nx, L.ptr(U), L.ptr(V),
               L.ptr(self.coef), L.ptr(self.hscale), C.addressof(self.dzeta), C.addressof(self.phys),
               L.ptr(Uo), L.ptr(Vo), L.ptr(mom_in), L.ptr(mom_out), C.addressof(adam),
               L.ptr(alpha_tab), int(nbitmax), L.ptr(partials), L.ptr(prev_partials), es_in, es_out,
               L.ptr(self.out), L.ptr(cost_hist), 1 if finalize_only else 0, L.stream_ptr())

    def finalize_pair(self, es=None, cost_hist=None, it=0, partials=None, out=None):
        """``finalize`` of a ``run_pass2`` call: iterations it and it+1."""
        out = self.out if out is None else out
        p = self.partials if partials is None else partials
        L.call("igm_energy_finalize_pair", L.ptr(p), self.Ny, self.Nx, L.ptr(out), es,
               L.ptr(cost_hist), it, L.stream_ptr())
        return out

    def finalize(self, out=None, es=None, cost_hist=None, it=0, partials=None):
        out = self.out if out is None else out
        p = self.partials if partials is None else partials
        L.call("igm_energy_finalize", L.ptr(p), self.B, self.Ny, self.Nx, L.ptr(out),
               es, L.ptr(cost_hist), it, L.stream_ptr())
        return out

    # -- convenience --------------------------------------------------------
    def energy(self, U, V=None):
        """Per-entry [mean shear, mean sliding, mean gravity, COST] (float64, device)."""
        self.run_pass(0, U, V)
        return self.finalize().clone()

    def energy_grad(self, U, V=None):
        """(energies (B,4) float64, gU, gV) -- or (energies, gY) for a Y input."""
        if V is None:
            gY = torch.empty_like(U)
            self.run_pass(1, U, None, gU=gY)
            return self.finalize().clone(), gY
        gU = torch.empty_like(U)
        gV = torch.empty_like(V)
        self.run_pass(1, U, V, gU=gU, gV=gV)
        return self.finalize().clone(), gU, gV


class EnergyFunction(torch.autograd.Function):
    """J(Y) = sum over batch entries of COST_b, with dJ/dY from the fused pass.

    Y: emulator output (B, Ny, Nx, 2Nz) float32.  Returns a float64 scalar.
    The reference sums the per-patch costs of a retraining epoch
    (emulate/emulate.py:222-226); each COST_b is already a mean over cells.
    """

    @staticmethod
    def forward(ctx, Y, op):
        Y = Y.contiguous()
        e, gY = op.energy_grad(Y)
        ctx.save_for_backward(gY)
        return e[:, 3].sum()

    @staticmethod
    def backward(ctx, g):
        (gY,) = ctx.saved_tensors
        return gY * g.to(gY.dtype), None
