"""GPU parity of the emulator's MFMA convolution kernels.

Each kernel is compared with torch-CPU float32 (or float64) on the SAME
bf16-representable operands, so only the accumulation order and the final
bf16 rounding of the output differ.  Tolerances (rel L2): 4e-3 for bf16
outputs (half an ulp of bf16 is 2^-9 = 2e-3), 1e-5 for fp32 outputs.
Whole-network checks compare with the oracle CNN (oracle/cnn.py) run with
bf16 rounding of weights and activations at the same places.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from igm_amd import _lib as L
from igm_amd import emulator as EM
from oracle import cnn as OC
from tests.helpers import rel, small_glacier

pytestmark = pytest.mark.gpu


def bf(x):
    return torch.as_tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float32)


def nchw(t):
    return t.permute(0, 3, 1, 2)


def nhwc(t):
    return t.permute(0, 2, 3, 1)


@pytest.mark.parametrize("B,H,W", [(1, 7, 5), (2, 33, 70), (1, 64, 64), (3, 17, 130),
                                   # more tiles than persistent blocks: each block walks a run
                                   (2, 255, 300)])
def test_conv3x3_fwd_and_dgrad(dev, B, H, W):
    print(f"case conv3x3 B={B} H={H} W={W}", flush=True)
    g = torch.Generator().manual_seed(H * W)
    x = bf(torch.randn((B, H, W, 32), generator=g))
    w = bf(torch.randn((3, 3, 32, 32), generator=g) * 0.1)  # HWIO
    b = torch.randn(32, generator=g) * 0.1
    cnn = EM.CNN(nz=10, nb_layers=2, device=dev)
    wts = cnn.get_weights()
    wts[1] = (w.numpy(), b.numpy())
    cnn.set_weights(wts)
    xd = x.to(torch.bfloat16).to(dev)
    out = torch.empty_like(xd)
    L.call("igm_conv3x3_fwd_bf16", L.ptr(xd), L.ptr(cnn.fwd[0]), L.ptr(cnn.bias(1)), None, B, H, W,
           L.ptr(out), L.stream_ptr())
    ref = F.leaky_relu(F.conv2d(nchw(x), w.permute(3, 2, 0, 1), b, padding=1), 0.01)
    assert rel(out.float().cpu(), nhwc(ref)) <= 4e-3
    # dgrad: dz_below = conv_transpose(dz) * leaky'(act_below)
    dz = bf(torch.randn((B, H, W, 32), generator=g))
    act_below = bf(torch.randn((B, H, W, 32), generator=g))
    dzd = dz.to(torch.bfloat16).to(dev)
    abd = act_below.to(torch.bfloat16).to(dev)
    outd = torch.empty_like(dzd)
    L.call("igm_conv3x3_dgrad_bf16", L.ptr(dzd), L.ptr(cnn.dgr[0]), L.ptr(abd), None, B, H, W, L.ptr(outd),
           L.stream_ptr())
    xi = nchw(act_below).clone().requires_grad_(True)
    y = F.conv2d(xi, w.permute(3, 2, 0, 1), None, padding=1)
    (gx,) = torch.autograd.grad(y, xi, nchw(dz))
    refd = nhwc(gx) * torch.where(act_below > 0, 1.0, 0.01)
    assert rel(outd.float().cpu(), refd) <= 4e-3


@pytest.mark.parametrize("B,H,W", [(1, 9, 11), (2, 40, 100)])
def test_wgrad_3x3_and_1x1(dev, B, H, W):
    g = torch.Generator().manual_seed(7 + H)
    a = bf(torch.randn((B, H, W, 32), generator=g))
    dz = bf(torch.randn((B, H, W, 32), generator=g))
    ws = torch.empty(L.load().igm_wgrad_workspace_floats(9, 64), dtype=torch.float32, device=dev)
    dW = torch.empty((3, 3, 32, 32), dtype=torch.float32, device=dev)
    db = torch.empty(32, dtype=torch.float32, device=dev)
    ad, dzd = a.to(torch.bfloat16).to(dev), dz.to(torch.bfloat16).to(dev)  # keep the buffers alive
    L.call("igm_conv_wgrad_bf16", 9, L.ptr(ad), L.ptr(dzd), None, B, H, W, 32, 64, L.ptr(ws), 0, L.ptr(dW),
           L.ptr(db), L.stream_ptr())
    wt = torch.zeros((32, 32, 3, 3), dtype=torch.float64, requires_grad=True)
    y = F.conv2d(nchw(a).double(), wt, None, padding=1)
    (gw,) = torch.autograd.grad(y, wt, nchw(dz).double())
    assert rel(dW.cpu(), gw.permute(2, 3, 1, 0)) <= 1e-5
    assert rel(db.cpu(), dz.double().sum(dim=(0, 1, 2))) <= 1e-5
    # 1x1 with cout = 20 (zero-padded bf16 dY)
    dY = torch.zeros_like(dz)
    dY[..., :20] = dz[..., :20]
    dW1 = torch.empty((1, 1, 32, 20), dtype=torch.float32, device=dev)
    db1 = torch.empty(20, dtype=torch.float32, device=dev)
    dYd = dY.to(torch.bfloat16).to(dev)
    L.call("igm_conv_wgrad_bf16", 1, L.ptr(ad), L.ptr(dYd), None, B, H, W, 20, 64, L.ptr(ws), 0, L.ptr(dW1),
           L.ptr(db1), L.stream_ptr())
    # accumulate: a second call adds the same gradient
    L.call("igm_conv_wgrad_bf16", 1, L.ptr(ad), L.ptr(dYd), None, B, H, W, 20, 64, L.ptr(ws), 1, L.ptr(dW1),
           L.ptr(db1), L.stream_ptr())
    ref = torch.einsum("bhwi,bhwo->io", a.double(), dY[..., :20].double())
    assert rel(dW1.cpu()[0, 0], 2 * ref) <= 1e-5
    assert rel(db1.cpu(), 2 * dY[..., :20].double().sum(dim=(0, 1, 2))) <= 1e-5


def test_conv_in_and_head(dev):
    B, H, W = 2, 37, 90
    gl = [small_glacier(H, W, seed=s) for s in range(B)]
    X = torch.stack([torch.stack([torch.as_tensor(gg[k]) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")], -1)
                     for gg in gl]).float()
    cnn = EM.CNN(nz=10, nb_layers=1, device=dev, seed=3)
    k0, b0 = cnn.get_weights()[0]
    Xd = X.to(dev).contiguous()
    fields = EM.nhwc_fields(Xd)
    out = torch.empty((B, H, W, 32), dtype=torch.bfloat16, device=dev)
    L.call("igm_conv_in_fwd", *fields[0], *fields[1], None, B, H, W, L.ptr(cnn.kernel(0)), L.ptr(cnn.bias(0)),
           L.ptr(out), L.stream_ptr())
    ref = nhwc(F.leaky_relu(F.conv2d(nchw(X).double(), torch.as_tensor(k0).double().permute(3, 2, 0, 1),
                                     torch.as_tensor(b0).double(), padding=1), 0.01))
    assert rel(out.float().cpu(), ref) <= 4e-3
    # head: 1x1 32 -> 20, fp32 out
    a = bf(torch.randn((B, H, W, 32)))
    Y = torch.empty((B, H, W, 20), dtype=torch.float32, device=dev)
    ad = a.to(torch.bfloat16).to(dev)
    L.call("igm_conv1x1_head_fwd", L.ptr(ad), L.ptr(cnn.head), L.ptr(cnn.bias(1)), None, B, H, W, 20, L.ptr(Y),
           L.stream_ptr())
    kh = bf(cnn.kernel(1).cpu())[0, 0]
    refY = a.double() @ kh.double() + cnn.bias(1).cpu().double()
    assert rel(Y.cpu(), refY) <= 1e-5


@pytest.mark.parametrize("cout,offset", [(20, 0), (20, 1), (18, 0), (7, 0), (32, 0), (32, 3)])
def test_head_store_paths_and_extents(dev, cout, offset):
    """The head writes 16-B stores when cout % 4 == 0 and Y is 16-B aligned,
    4-B stores otherwise (odd cout, or Y offset by `offset` floats); pixels
    outside a sample's extent are zero and channels beyond cout untouched."""
    B, H, W = 3, 21, 45
    g = torch.Generator().manual_seed(cout * 10 + offset)
    a = bf(torch.randn((B, H, W, 32), generator=g))
    wt = bf(torch.randn((32, 32), generator=g) * 0.2)
    wt[cout:] = 0.0  # rows beyond cout are the zero padding of the packed layout
    bias = torch.randn(cout, generator=g)
    ext = torch.tensor([[H, W], [13, 30], [1, W]], dtype=torch.int32)
    n = B * H * W * cout
    buf = torch.full((n + offset + 8,), 7.0, dtype=torch.float32, device=dev)
    Y = buf[offset:offset + n]
    ad = a.to(torch.bfloat16).to(dev)
    wd = wt.to(torch.bfloat16).to(dev)
    bd = bias.to(dev)
    ed = ext.to(dev)
    L.call("igm_conv1x1_head_fwd", L.ptr(ad), L.ptr(wd), L.ptr(bd), L.ptr(ed), B, H, W, cout, L.ptr(Y),
           L.stream_ptr())
    torch.cuda.synchronize()
    refY = (a.double() @ wt[:cout].double().T + bias.double()).reshape(B, H, W, cout)
    for b in range(B):
        ny, nx = int(ext[b, 0]), int(ext[b, 1])
        keep = torch.zeros((H, W), dtype=torch.bool)
        keep[:ny, :nx] = True
        refY[b][~keep] = 0.0
    out = buf.cpu()
    assert rel(out[offset:offset + n].reshape(B, H, W, cout), refY) <= 1e-5
    # outside the padded region nothing is written
    assert torch.all(out[:offset] == 7.0) and torch.all(out[offset + n:] == 7.0)
    # the zeroed pixels are exactly zero
    assert torch.all(out[offset:offset + n].reshape(B, H, W, cout)[1, 13:] == 0.0)


@pytest.mark.parametrize("cout,offset", [(20, 0), (20, 1), (18, 0), (10, 0), (10, 3), (7, 0), (32, 2)])
def test_head_bwd_load_paths_and_extents(dev, cout, offset):
    """head_bwd reads dY with 16-B loads when cout % 4 == 0 and dY is 16-B
    aligned, 4-B loads otherwise (cout 18, 10, 7, or dY offset by `offset`
    floats).  dz16 = bf16((dY W^T) * LeakyReLU'(a16)) and dY_bf16 = bf16(dY)
    zero-padded to 32 channels; both are zero outside a sample's extent."""
    print(f"case head_bwd cout={cout} offset={offset}", flush=True)
    B, H, W = 3, 19, 47
    g = torch.Generator().manual_seed(100 + cout * 10 + offset)
    dY = torch.randn((B, H, W, cout), generator=g)
    a = bf(torch.randn((B, H, W, 32), generator=g))
    w = torch.randn((32, cout), generator=g) * 0.2  # HWIO [32 in][cout]
    ext = torch.tensor([[H, W], [11, 30], [H, 1]], dtype=torch.int32)
    n = B * H * W * cout
    buf = torch.zeros((n + offset + 4,), dtype=torch.float32, device=dev)
    dYd = buf[offset:offset + n]
    dYd.copy_(dY.reshape(-1).to(dev))
    ad = a.to(torch.bfloat16).to(dev)
    wd = w.to(dev).contiguous()
    dz = torch.full((B, H, W, 32), 3.0, dtype=torch.bfloat16, device=dev)
    dYb = torch.full((B, H, W, 32), 3.0, dtype=torch.bfloat16, device=dev)
    ed = ext.to(dev)
    L.call("igm_conv1x1_head_bwd", L.ptr(dYd), L.ptr(ad), L.ptr(wd), L.ptr(ed), B, H, W, cout,
           L.ptr(dz), L.ptr(dYb), L.stream_ptr())
    torch.cuda.synchronize()
    slope = torch.where(a > 0, 1.0, 0.01).double()
    ref_dz = (dY.double() @ w.double().T) * slope
    ref_dYb = torch.zeros((B, H, W, 32))
    ref_dYb[..., :cout] = dY.to(torch.bfloat16).float()
    for b in range(B):
        keep = torch.zeros((H, W), dtype=torch.bool)
        keep[:int(ext[b, 0]), :int(ext[b, 1])] = True
        ref_dz[b][~keep] = 0.0
        ref_dYb[b][~keep] = 0.0
    got_dz = dz.float().cpu()
    assert rel(got_dz, ref_dz) <= 4e-3
    # the bf16 copy of dY is the round-to-nearest-even of each value, exactly
    assert torch.equal(dYb.float().cpu(), ref_dYb)
    assert torch.all(got_dz[1, 11:] == 0.0) and torch.all(got_dz[2, :, 1:] == 0.0)


def test_conv_in_wgrad(dev):
    B, H, W = 2, 30, 300
    g = torch.Generator().manual_seed(5)
    X = torch.randn((B, H, W, 5), generator=g)
    dz = bf(torch.randn((B, H, W, 32), generator=g))
    ws = torch.empty(L.load().igm_wgrad_workspace_floats(0, 64), dtype=torch.float32, device=dev)
    dW = torch.empty((3, 3, 5, 32), dtype=torch.float32, device=dev)
    db = torch.empty(32, dtype=torch.float32, device=dev)
    Xd = X.to(dev).contiguous()
    fields = EM.nhwc_fields(Xd)
    dzd = dz.to(torch.bfloat16).to(dev)
    L.call("igm_conv_in_wgrad", *fields[0], *fields[1], None, B, H, W, L.ptr(dzd), 64, L.ptr(ws), 0, L.ptr(dW),
           L.ptr(db), L.stream_ptr())
    wt = torch.zeros((32, 5, 3, 3), dtype=torch.float64, requires_grad=True)
    y = F.conv2d(nchw(X).double(), wt, None, padding=1)
    (gw,) = torch.autograd.grad(y, wt, nchw(dz).double())
    assert rel(dW.cpu(), gw.permute(2, 3, 1, 0)) <= 1e-5
    assert rel(db.cpu(), dz.double().sum(dim=(0, 1, 2))) <= 1e-5


def oracle_forward_bf16(X, weights):
    """Oracle CNN with the build's rounding points: bf16 weights for layers
    >= 2 and the head, bf16 activations after every hidden layer."""
    h = nchw(torch.as_tensor(X, dtype=torch.float64))
    for li, (k, b) in enumerate(weights):
        kt = torch.as_tensor(k, dtype=torch.float64)
        if li > 0:
            kt = bf(kt).double()
        h = F.conv2d(h, kt.permute(3, 2, 0, 1), torch.as_tensor(b, dtype=torch.float64), padding=kt.shape[0] // 2)
        if li < len(weights) - 1:
            h = bf(F.leaky_relu(h, 0.01)).double()
    return nhwc(h)


def test_network_forward_pretrained(dev):
    """IGM's shipped emulator weights on a glacier: full 17-layer forward."""
    weights, _ = EM.load_weights(EM.os.path.join(EM.EMULATOR_DIR, "pinnbp_10_4_cnn_16_32_2_1_a.safetensors"))
    cnn = EM.CNN(nz=10, nb_layers=16, device=dev, weights=weights)
    g = small_glacier(48, 80, dx=100.0, seed=1)
    X = torch.stack([torch.as_tensor(g[k]) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")], -1)[None].float()
    Y = cnn.forward(EM.nhwc_fields(X.to(dev).contiguous()))
    ref = oracle_forward_bf16(X, weights)
    assert rel(Y.cpu(), ref) <= 1e-2
    # and the plain fp32 oracle (what Keras computes) within the bf16 budget
    # IGM's weights were trained in fp32 on raw inputs: under bf16 activations
    # they lose accuracy (recorded, not asserted: the network the build ships
    # is trained through these bf16 kernels)
    ref32 = OC.forward(X, weights, dtype=torch.float64)
    print("IGM pretrained net under bf16: rel L2 vs fp64 oracle", rel(Y.cpu(), ref32))


def test_network_backward_vs_autograd(dev):
    """Parameter gradient of sum(dY * Y) vs torch autograd of the oracle CNN
    (float64, bf16 rounding simulated with a straight-through estimator)."""
    nl = 4
    cnn = EM.CNN(nz=10, nb_layers=nl, device=dev, seed=11)
    weights = cnn.get_weights()
    g = small_glacier(24, 70, dx=50.0, seed=4)
    X = torch.stack([torch.as_tensor(g[k]) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")], -1)[None].float()
    Xd = X.to(dev).contiguous()
    fields = EM.nhwc_fields(Xd)
    Y, acts = cnn.forward(fields, keep=True)
    dY = torch.randn(Y.shape, generator=torch.Generator().manual_seed(2)).float()
    dX = torch.empty(tuple(Xd.shape), dtype=torch.float32, device=dev)
    cnn.backward(dY.to(dev), acts, fields, input_grad=dX)
    grads = [(cnn.kernel(i, cnn.g).cpu(), cnn.bias(i, cnn.g).cpu()) for i in range(nl + 1)]

    def ste(x):  # bf16 rounding forward, identity backward
        return x + (bf(x.detach()).double() - x.detach())

    ps = []
    Xr = X.double().requires_grad_(True)
    h = nchw(Xr)
    for li, (k, b) in enumerate(weights):
        kt = torch.as_tensor(k, dtype=torch.float64).requires_grad_(True)
        bt = torch.as_tensor(b, dtype=torch.float64).requires_grad_(True)
        ps += [kt, bt]
        kk = ste(kt) if li > 0 else kt
        h = F.conv2d(h, kk.permute(3, 2, 0, 1), bt, padding=kt.shape[0] // 2)
        if li < nl:
            h = ste(F.leaky_relu(h, 0.01))
    loss = (nhwc(h) * dY.double()).sum()
    gr = torch.autograd.grad(loss, ps + [Xr])
    for li in range(nl + 1):
        ek, eb = rel(grads[li][0], gr[2 * li]), rel(grads[li][1], gr[2 * li + 1])
        print(f"layer {li}: kernel grad rel {ek:.2e}, bias grad rel {eb:.2e}")
        assert ek <= 3e-2 and eb <= 3e-2
    # gradient w.r.t. the input fields (data assimilation through the emulator)
    ex = rel(dX.cpu(), gr[-1])
    print(f"input grad rel {ex:.2e}")
    assert ex <= 3e-2


def test_padded_batch_equals_single(dev):
    cnn = EM.CNN(nz=10, nb_layers=3, device=dev, seed=5)
    sizes = [(20, 33), (40, 64), (9, 100)]
    H, W = 40, 100
    X = torch.zeros((3, H, W, 5))
    singles = []
    for b, (ny, nx) in enumerate(sizes):
        g = small_glacier(ny, nx, seed=b)
        xb = torch.stack([torch.as_tensor(g[k]) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")], -1).float()
        X[b, :ny, :nx] = xb
        singles.append(xb)
    X[0, 25:, :, :] = 7.0  # garbage outside the extents must not leak in
    ext = torch.tensor(sizes, dtype=torch.int32, device=dev)
    Y = cnn.forward(EM.nhwc_fields(X.to(dev).contiguous()), extents=ext).cpu()
    for b, (ny, nx) in enumerate(sizes):
        Yb = cnn.forward(EM.nhwc_fields(singles[b][None].to(dev).contiguous())).cpu()
        assert torch.equal(Y[b, :ny, :nx], Yb[0])
        assert float(Y[b, ny:].abs().sum()) == 0.0 and float(Y[b, :, nx:].abs().sum()) == 0.0


def test_padded_batch_many_tiles(dev):
    """Padded batch whose persistent blocks walk runs mixing valid and
    padding-only tiles: every entry equals its own unpadded forward."""
    cnn = EM.CNN(nz=10, nb_layers=3, device=dev, seed=9)
    sizes = [(300, 200), (150, 260), (37, 61)]
    H, W = 300, 260
    X = torch.zeros((3, H, W, 5))
    singles = []
    for b, (ny, nx) in enumerate(sizes):
        g = small_glacier(ny, nx, seed=30 + b)
        xb = torch.stack([torch.as_tensor(g[k]) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")], -1).float()
        X[b, :ny, :nx] = xb
        singles.append(xb)
    ext = torch.tensor(sizes, dtype=torch.int32, device=dev)
    Y = cnn.forward(EM.nhwc_fields(X.to(dev).contiguous()), extents=ext).cpu()
    for b, (ny, nx) in enumerate(sizes):
        Yb = cnn.forward(EM.nhwc_fields(singles[b][None].to(dev).contiguous())).cpu()
        assert torch.equal(Y[b, :ny, :nx], Yb[0])
        assert float(Y[b, ny:].abs().sum()) == 0.0 and float(Y[b, :, nx:].abs().sum()) == 0.0


def test_adam_clip(dev):
    cnn = EM.CNN(nz=10, nb_layers=2, device=dev, seed=1)
    w0 = cnn.w.cpu().double().clone()
    g = torch.randn(cnn.n, generator=torch.Generator().manual_seed(3)).float()
    cnn.g.copy_(g.to(dev))
    cnn.adam_step(lr=1e-3, clipnorm=1.0)
    # oracle: per-tensor clip_by_norm then legacy Adam at t = 1
    w = w0.clone()
    for i in range(len(cnn.shapes)):
        a, e = cnn.offs_host[i], cnn.offs_host[i + 1]
        gi = g[a:e].double()
        n = float(torch.linalg.vector_norm(gi))
        gi = gi * 1.0 / max(n, 1.0)
        m = 0.1 * gi
        v = 0.001 * gi * gi
        alpha = 1e-3 * np.sqrt(1 - 0.999) / (1 - 0.9)
        w[a:e] -= (m * alpha) / (torch.sqrt(v) + 1e-7)
    assert rel(cnn.w.cpu(), w) <= 1e-6


def test_bucketed_gradient_equals_padded(dev):
    """Sub-batches with accumulated gradients give the gradient of the whole
    zero-padded batch (the padding is masked, the sums only reorder)."""
    from igm_amd.energy import EnergyOperator
    from oracle.energy import DEFAULT_PHYSICS
    from tests.helpers import dzeta_for
    cnn = EM.CNN(nz=4, nb_layers=3, device=dev, seed=7)
    dz = dzeta_for(4)
    sizes = [(24, 30), (48, 64), (16, 90), (40, 40)]
    parts = []
    for b, (ny, nx) in enumerate(sizes):
        g = small_glacier(ny, nx, seed=10 + b)
        parts.append(torch.stack([torch.as_tensor(g[k]) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")],
                                 -1).float().to(dev))

    def pack(ps):
        H = max(p.shape[0] for p in ps)
        W = max(p.shape[1] for p in ps)
        X = torch.zeros((len(ps), H, W, 5), device=dev)
        for b, p in enumerate(ps):
            X[b, :p.shape[0], :p.shape[1]] = p
        ext = torch.tensor([list(p.shape[:2]) for p in ps], dtype=torch.int32, device=dev)
        return X, ext

    def grad(groups, deferred=False):
        keep = []
        for k, ps in enumerate(groups):
            X, ext = pack(ps)
            op = EnergyOperator(dz, DEFAULT_PHYSICS, X=X, extents=ext)
            fields = EM.nhwc_fields(X)
            Y, acts = cnn.forward(fields, extents=ext, keep=True)
            e, gY = op.energy_grad(Y)
            cnn.backward(gY, acts, fields, extents=ext, accumulate=k > 0, deferred=deferred)
            keep.append((X, ext, fields, Y, acts, gY))
        if deferred:
            cnn.reduce_deferred()
        torch.cuda.synchronize()
        return cnn.g.cpu().clone()

    g_all = grad([parts])
    g_bkt = grad([[parts[0], parts[3]], [parts[1], parts[2]]])
    assert float(g_all.abs().max()) > 0
    assert rel(g_bkt, g_all) <= 1e-5
    # deferred: the sub-batches add to per-block partials, one reduction per
    # layer; a second batch overwrites the consumed partials (same bits)
    g_def = grad([[parts[0], parts[3]], [parts[1], parts[2]]], deferred=True)
    assert rel(g_def, g_all) <= 1e-5
    g_def2 = grad([[parts[0], parts[3]], [parts[1], parts[2]]], deferred=True)
    assert torch.equal(g_def2, g_def)
    g_one = grad([parts], deferred=True)  # one sub-batch: the plain backward's bits
    assert torch.equal(g_one, g_all)
