"""GPU parity of the fused energy / adjoint / Adam pass against the CPU oracle.

Tolerances (written per test): the float32 kernel vs the float64 oracle is
held to the float32 rounding of a ~10-term staggered stencil (COST rel err
<= 2e-6, gradient rel L2 <= 2e-5); the float64 kernel vs the float64 oracle
to 1e-11 (only summation order differs).
"""
import numpy as np
import pytest
import torch

from igm_amd.energy import EnergyOperator
from oracle import energy as E
from oracle.solver import LegacyAdam
from tests.helpers import dzeta_for, rel, small_glacier, smooth_velocity, to_dev

pytestmark = pytest.mark.gpu

PHYS = E.DEFAULT_PHYSICS


def oracle_eval(g, U, V, nz, dtype=torch.float64):
    fin = E.fieldin_tensors(g["thk"], g["usurf"], g["arrhenius"], g["slidingco"], g["dX"], dtype)
    c, means, gU, gV = E.energy_and_grad(torch.as_tensor(U, dtype=dtype), torch.as_tensor(V, dtype=dtype),
                                         fin, torch.as_tensor(dzeta_for(nz), dtype=dtype))
    return float(c), [float(m) for m in means], gU, gV


def gpu_op(g, nz, dev, dtype):
    fields = {k: to_dev(g[k], dev, dtype) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")}
    return EnergyOperator(dzeta_for(nz), PHYS, fields=fields, dtype=dtype)


SHAPES = [(2, 2), (3, 5), (9, 8), (17, 61), (24, 62), (31, 63), (40, 64), (33, 125), (50, 124), (64, 200)]


@pytest.mark.parametrize("ny,nx", SHAPES)
def test_energy_and_gradient_f32_shapes(dev, ny, nx):
    nz = 10
    print(f"case energy f32 ny={ny} nx={nx}", flush=True)
    g = small_glacier(ny, nx, seed=ny * 1000 + nx)
    U, V = smooth_velocity(nz, ny, nx, seed=nx)
    c_ref, m_ref, gU_ref, gV_ref = oracle_eval(g, U, V, nz)
    op = gpu_op(g, nz, dev, torch.float32)
    e, gU, gV = op.energy_grad(to_dev(U, dev), to_dev(V, dev))
    e = e.cpu().numpy()[0]
    for k in range(3):
        assert abs(e[k] - m_ref[k]) <= 2e-6 * abs(c_ref) + 1e-30, (k, e[k], m_ref[k])
    assert abs(e[3] - c_ref) <= 2e-6 * abs(c_ref)
    g_ref = torch.stack([gU_ref, gV_ref])
    assert rel(torch.stack([gU.cpu(), gV.cpu()]), g_ref) <= 2e-5


@pytest.mark.parametrize("nz", [2, 3, 5, 10, 20])
def test_energy_and_gradient_f64_nz(dev, nz):
    ny, nx = 21, 70
    print(f"case energy f64 nz={nz}", flush=True)
    g = small_glacier(ny, nx, seed=nz)
    U, V = smooth_velocity(nz, ny, nx, seed=nz)
    c_ref, m_ref, gU_ref, gV_ref = oracle_eval(g, U, V, nz)
    op = gpu_op(g, nz, dev, torch.float64)
    e, gU, gV = op.energy_grad(to_dev(U, dev, torch.float64), to_dev(V, dev, torch.float64))
    e = e.cpu().numpy()[0]
    assert abs(e[3] - c_ref) <= 1e-11 * abs(c_ref)
    assert rel(torch.stack([gU.cpu(), gV.cpu()]), torch.stack([gU_ref, gV_ref])) <= 1e-11


def test_gradient_vs_central_differences_f64(dev):
    """Brute force: central differences of the GPU float64 energy at random entries."""
    nz, ny, nx = 5, 12, 70
    g = small_glacier(ny, nx, seed=5)
    U, V = smooth_velocity(nz, ny, nx, seed=6)
    op = gpu_op(g, nz, dev, torch.float64)
    Ud, Vd = to_dev(U, dev, torch.float64), to_dev(V, dev, torch.float64)
    _, gU, gV = op.energy_grad(Ud, Vd)
    rng = np.random.default_rng(0)
    for _ in range(12):
        which = rng.integers(2)
        k, j, i = rng.integers(nz), rng.integers(ny), rng.integers(nx)
        h = 1e-3
        T = (Ud if which == 0 else Vd)
        orig = float(T[k, j, i])
        T[k, j, i] = orig + h
        ep = float(op.energy(Ud, Vd)[0, 3])
        T[k, j, i] = orig - h
        em = float(op.energy(Ud, Vd)[0, 3])
        T[k, j, i] = orig
        fd = (ep - em) / (2 * h)
        an = float((gU if which == 0 else gV)[k, j, i])
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-9) + 1e-12, (which, k, j, i, fd, an)


def test_layout_nhwc_equals_planar(dev):
    nz, ny, nx = 10, 30, 90
    g = small_glacier(ny, nx, seed=3)
    U, V = smooth_velocity(nz, ny, nx, seed=3)
    op = gpu_op(g, nz, dev, torch.float32)
    e1, gU, gV = op.energy_grad(to_dev(U, dev), to_dev(V, dev))
    Y = torch.cat([to_dev(U, dev).permute(1, 2, 0), to_dev(V, dev).permute(1, 2, 0)], -1)[None].contiguous()
    e2, gY = op.energy_grad(Y)
    assert torch.equal(e1, e2)
    assert torch.equal(gY[0, ..., :nz].permute(2, 0, 1), gU)
    assert torch.equal(gY[0, ..., nz:].permute(2, 0, 1), gV)


def test_layout_nhwc_unaligned_equals_aligned(dev):
    """A Y view that is not 16-byte aligned takes the one-level-per-access
    path; a 16-byte aligned Y moves whole nodes as 16-B accesses.  Same
    arithmetic, so the same energies and gradient bit for bit (Nz 10 and the
    odd Nz 5, which always takes the one-level path)."""
    for nz in (10, 5):
        ny, nx = 37, 70
        g = small_glacier(ny, nx, seed=nz)
        U, V = smooth_velocity(nz, ny, nx, seed=nz + 1)
        op = gpu_op(g, nz, dev, torch.float32)
        Y = torch.cat([to_dev(U, dev).permute(1, 2, 0), to_dev(V, dev).permute(1, 2, 0)], -1)[None].contiguous()
        assert Y.data_ptr() % 16 == 0
        buf = torch.empty(Y.numel() + 1, dtype=torch.float32, device=dev)
        Yu = buf[1:].view(Y.shape)
        Yu.copy_(Y)
        assert Yu.data_ptr() % 16 == 4
        e1, g1 = op.energy_grad(Y)
        e2, g2 = op.energy_grad(Yu)
        assert torch.equal(e1, e2), nz
        assert torch.equal(g1, g2), nz


def test_batch_with_extents_equals_single(dev):
    """Patches of uneven sizes packed in one padded batch: masking by extents."""
    nz = 10
    sizes = [(20, 33), (40, 64), (9, 100)]
    H, W = 40, 100
    B = len(sizes)
    X = torch.zeros((B, H, W, 5), dtype=torch.float32)
    Y = torch.randn((B, H, W, 2 * nz), dtype=torch.float32) * 50  # garbage outside the patches
    singles = []
    for b, (ny, nx) in enumerate(sizes):
        g = small_glacier(ny, nx, seed=b + 11)
        for f, k in enumerate(("thk", "usurf", "arrhenius", "slidingco", "dX")):
            X[b, :ny, :nx, f] = torch.as_tensor(g[k])
        U, V = smooth_velocity(nz, ny, nx, seed=b)
        Y[b, :ny, :nx, :nz] = torch.as_tensor(U, dtype=torch.float32).permute(1, 2, 0)
        Y[b, :ny, :nx, nz:] = torch.as_tensor(V, dtype=torch.float32).permute(1, 2, 0)
        singles.append((g, U, V))
    ext = torch.tensor(sizes, dtype=torch.int32, device=dev)
    op = EnergyOperator(dzeta_for(nz), PHYS, X=X.to(dev), extents=ext)
    e, gY = op.energy_grad(Y.to(dev))
    e = e.cpu(); gY = gY.cpu()
    for b, (g, U, V) in enumerate(singles):
        ny, nx = sizes[b]
        opb = gpu_op(g, nz, dev, torch.float32)
        eb, gU, gV = opb.energy_grad(to_dev(U, dev), to_dev(V, dev))
        assert abs(float(e[b, 3]) - float(eb[0, 3])) <= 1e-6 * abs(float(eb[0, 3]))
        assert rel(gY[b, :ny, :nx, :nz].permute(2, 0, 1), gU.cpu()) <= 1e-6
        assert float(gY[b, ny:, :, :].abs().max() if ny < H else 0) == 0.0
        assert float(gY[b, :, nx:, :].abs().max() if nx < W else 0) == 0.0


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_adam_step(dev, dtype):
    """mode 2: energy + adjoint + one legacy-Adam step vs the oracle's step."""
    nz, ny, nx = 10, 25, 130
    g = small_glacier(ny, nx, seed=21)
    U, V = smooth_velocity(nz, ny, nx, seed=21)
    op = gpu_op(g, nz, dev, dtype)
    from igm_amd import _lib as L
    from igm_amd.solver import LegacyAdamState
    st = LegacyAdamState(lr=1.0)
    Ud, Vd = to_dev(U, dev, dtype), to_dev(V, dev, dtype)
    Uo, Vo = torch.empty_like(Ud), torch.empty_like(Vd)
    m = [torch.zeros_like(Ud) for _ in range(4)]
    rng = np.random.default_rng(1)
    m[1] += torch.as_tensor(rng.uniform(0, 1e-3, Ud.shape), dtype=dtype, device=dev)  # non-zero v
    m[3] += torch.as_tensor(rng.uniform(0, 1e-3, Ud.shape), dtype=dtype, device=dev)
    m_host = [t.cpu().double().clone() for t in m]
    adam = L.IgmAdam(alpha=st.alpha(3, dtype), beta1=0.9, beta2=0.999, epsilon=1e-7)
    mom = torch.stack(m, -1).contiguous()  # (m_U, v_U, m_V, v_V) per node-level
    op.run_pass(2, Ud, Vd, Uo=Uo, Vo=Vo, moments=mom, adam=adam)
    m = [mom[..., i] for i in range(4)]
    # oracle
    _, _, gU, gV = oracle_eval(g, U, V, nz)
    opt = LegacyAdam(lr=1.0, iterations=2)
    x = [torch.as_tensor(U, dtype=torch.float64).clone(), torch.as_tensor(V, dtype=torch.float64).clone()]
    opt.apply(x, [gU, gV], [m_host[0], m_host[2]], [m_host[1], m_host[3]])
    tol = 1e-5 if dtype == torch.float32 else 1e-12
    assert rel(Uo.cpu(), x[0]) <= tol and rel(Vo.cpu(), x[1]) <= tol
    assert rel(m[0].cpu(), m_host[0]) <= 20 * tol
    assert rel(m[3].cpu(), m_host[3]) <= 20 * tol


def test_zero_velocity_closed_form(dev):
    """U=V=0: shear and gravity vanish; sliding = mean(C_s (regu_w^2)^(2/3) / (4/3))."""
    nz, ny, nx = 10, 33, 77
    g = small_glacier(ny, nx, seed=2, C=0.05)
    op = gpu_op(g, nz, dev, torch.float64)
    z = torch.zeros((nz, ny, nx), dtype=torch.float64, device=dev)
    e = op.energy(z, z).cpu().numpy()[0]
    assert e[0] == 0.0 and e[2] == 0.0
    # the fields are float32: slidingco is float32(0.05)
    expect = float(np.float32(0.05)) * (1e-20) ** (2.0 / 3.0) / (4.0 / 3.0)
    assert abs(e[1] - expect) <= 1e-12 * expect


def test_deterministic_bitwise(dev):
    nz, ny, nx = 10, 80, 150
    g = small_glacier(ny, nx, seed=8)
    U, V = smooth_velocity(nz, ny, nx, seed=8)
    op = gpu_op(g, nz, dev, torch.float32)
    Ud, Vd = to_dev(U, dev), to_dev(V, dev)
    r1 = op.energy_grad(Ud, Vd)
    r2 = op.energy_grad(Ud, Vd)
    for a, b in zip(r1, r2):
        assert torch.equal(a, b)


def floating_glacier(ny, nx, seed):
    """small_glacier with a sliding field that is zero on one block, negative
    on another and mixed in sign along their edges, so that the cell average
    of slidingco is <= 0 (floating-ice damping, cost_shear.py:85-90) on whole
    blocks and on cells with some positive corners."""
    g = small_glacier(ny, nx, seed=seed)
    rng = np.random.default_rng(seed)
    C = np.full((ny, nx), 0.0464, np.float32)
    C[ny // 4:ny // 2, nx // 5:nx // 2] = 0.0
    C[ny // 2:, nx // 2:(3 * nx) // 4] = -0.03
    # scattered single negative corners (some outweigh their three positive neighbours)
    for _ in range(max(1, ny * nx // 40)):
        C[rng.integers(ny), rng.integers(nx)] = rng.choice([-0.2, -0.01, 0.0])
    g["slidingco"] = C
    return g


@pytest.mark.parametrize("ny,nx", [(17, 61), (40, 64), (50, 124)])
@pytest.mark.parametrize("dtype,ctol,gtol", [(torch.float32, 2e-6, 2e-5), (torch.float64, 1e-11, 1e-11)])
def test_floating_damping_vs_oracle(dev, ny, nx, dtype, ctol, gtol):
    nz = 10
    print(f"case floating damping {dtype} ny={ny} nx={nx}", flush=True)
    g = floating_glacier(ny, nx, seed=ny + nx)
    C = g["slidingco"]
    cs = (C[:-1, :-1] + C[:-1, 1:] + C[1:, :-1] + C[1:, 1:]) / 4
    assert (cs <= 0).sum() >= 20 and (cs > 0).sum() >= 20
    U, V = smooth_velocity(nz, ny, nx, seed=ny)
    c_ref, m_ref, gU_ref, gV_ref = oracle_eval(g, U, V, nz)
    op = gpu_op(g, nz, dev, dtype)
    e, gU, gV = op.energy_grad(to_dev(U, dev, dtype), to_dev(V, dev, dtype))
    e = e.cpu().numpy()[0]
    scale = sum(abs(m) for m in m_ref)
    for k in range(3):
        assert abs(e[k] - m_ref[k]) <= ctol * scale, (k, e[k], m_ref[k])
    assert abs(e[3] - c_ref) <= ctol * scale
    assert rel(torch.stack([gU.cpu(), gV.cpu()]), torch.stack([gU_ref, gV_ref])) <= gtol
    # the damped cells change the energy: the same field with slidingco
    # clipped to > 0 has a different shear energy
    g2 = dict(g)
    g2["slidingco"] = np.maximum(C, np.float32(1e-3))
    _, m2, _, _ = oracle_eval(g2, U, V, nz)
    assert abs(m2[0] - m_ref[0]) > 100 * ctol * scale
