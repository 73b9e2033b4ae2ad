"""igm_energy_eval_f32_noalloc, the entry point that allocates nothing,
called through the C-ABI: the same energies and gradient, bit for bit, as
the workspace entry points igm_energy_pass_f32 + igm_energy_finalize on the
same inputs -- on two calls in a row, the second with a smaller input (the
static scratch keeps the first call's partial sums) -- and a clean error
for a batch whose partial sums exceed the static scratch."""
import ctypes as C

import pytest
import torch

from igm_amd import _lib as L
from igm_amd.energy import EnergyOperator, velocity_strides
from oracle.energy import DEFAULT_PHYSICS
from tests.helpers import dzeta_for, small_glacier, smooth_velocity, to_dev

pytestmark = pytest.mark.gpu
KEYS = ("thk", "usurf", "arrhenius", "slidingco", "dX")


def workspace_eval(op, mode, U, V=None):
    """The existing entry points: pass with the operator's own workspace, then finalize."""
    if V is None:
        g = torch.zeros_like(U) if mode == 1 else None
        op.run_pass(mode, U, None, gU=g)
        return op.finalize().clone(), g
    gU = torch.zeros_like(U) if mode == 1 else None
    gV = torch.zeros_like(V) if mode == 1 else None
    op.run_pass(mode, U, V, gU=gU, gV=gV)
    return op.finalize().clone(), (gU, gV)


def noalloc_eval(op, mode, U, V=None):
    out = torch.full((op.B, 4), float("nan"), dtype=torch.float64, device=U.device)
    up, vp, bs, ks, rs, cs = velocity_strides(U, V, op.nz)
    if V is None:
        g = torch.zeros_like(U) if mode == 1 else None
        gu, gv = (velocity_strides(g, None, op.nz)[:2] if g is not None else (None, None))
        grads = g
    else:
        gU = torch.zeros_like(U) if mode == 1 else None
        gV = torch.zeros_like(V) if mode == 1 else None
        gu, gv = L.ptr(gU), L.ptr(gV)
        grads = (gU, gV)
    L.call("igm_energy_eval_f32_noalloc", mode, op.nz, op.B, op.Ny, op.Nx, up, vp, bs, ks, rs, cs,
           L.ptr(op.coef), L.ptr(op.hscale), C.addressof(op.dzeta), C.addressof(op.phys), gu, gv,
           L.ptr(out), L.stream_ptr())
    return out, grads


def same(a, b):
    if isinstance(a, tuple):
        return all(same(x, y) for x, y in zip(a, b))
    if a is None:
        return b is None
    return torch.equal(a, b)


def test_noalloc_equals_workspace_entry_points(dev):
    nz = 10
    # call 1: a padded batch of three uneven entries, emulator (Y) layout
    print("case noalloc call 1: batch of 3, 150 x 260, Y layout", flush=True)
    sizes = [(150, 260), (97, 130), (40, 201)]
    H, W = 150, 260
    X = torch.zeros((3, H, W, 5))
    Y = torch.zeros((3, H, W, 2 * nz))
    for b, (ny, nx) in enumerate(sizes):
        g = small_glacier(ny, nx, seed=40 + b)
        for f, k in enumerate(KEYS):
            X[b, :ny, :nx, f] = torch.as_tensor(g[k])
        U, V = smooth_velocity(nz, ny, nx, seed=b)
        Y[b, :ny, :nx, :nz] = torch.as_tensor(U, dtype=torch.float32).permute(1, 2, 0)
        Y[b, :ny, :nx, nz:] = torch.as_tensor(V, dtype=torch.float32).permute(1, 2, 0)
    ext = torch.tensor(sizes, dtype=torch.int32, device=dev)
    op = EnergyOperator(dzeta_for(nz), DEFAULT_PHYSICS, X=X.to(dev).contiguous(), extents=ext)
    Yd = Y.to(dev).contiguous()
    for mode in (0, 1):
        e_ws, g_ws = workspace_eval(op, mode, Yd)
        e_na, g_na = noalloc_eval(op, mode, Yd)
        assert torch.equal(e_ws, e_na), (mode, e_ws, e_na)
        assert same(g_ws, g_na), mode
    assert float(e_na[:, 3].abs().min()) > 0
    # call 2: one smaller planar field (fewer blocks than call 1 left behind)
    print("case noalloc call 2: one 40 x 50 field, planar", flush=True)
    g = small_glacier(40, 50, seed=7)
    op2 = EnergyOperator(dzeta_for(nz), DEFAULT_PHYSICS, fields={k: to_dev(g[k], dev) for k in KEYS})
    U, V = smooth_velocity(nz, 40, 50, seed=9)
    Ud, Vd = to_dev(U, dev), to_dev(V, dev)
    for mode in (1, 0):
        e_ws, g_ws = workspace_eval(op2, mode, Ud, Vd)
        e_na, g_na = noalloc_eval(op2, mode, Ud, Vd)
        assert torch.equal(e_ws, e_na), (mode, e_ws, e_na)
        assert same(g_ws, g_na), mode
    torch.cuda.synchronize()


def test_noalloc_rejects_a_batch_beyond_the_scratch(dev):
    """30,000 entries of 2 x 64 need 90,000 partial sums: rejected on the
    host with a clear message, before anything is launched."""
    t = torch.zeros(16, dtype=torch.float32, device=dev)
    o = torch.zeros(4, dtype=torch.float64, device=dev)
    dz = (C.c_double * 9)(*dzeta_for(10))
    from igm_amd.energy import physics_struct
    ph = physics_struct(DEFAULT_PHYSICS)
    with pytest.raises(L.IgmError, match="static scratch"):
        L.call("igm_energy_eval_f32_noalloc", 0, 10, 30000, 2, 64, L.ptr(t), L.ptr(t), 2 * 64 * 10, 2 * 64,
               64, 1, L.ptr(t), L.ptr(t), C.addressof(dz), C.addressof(ph), None, None, L.ptr(o),
               L.stream_ptr())
