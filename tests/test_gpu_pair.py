"""The two-iteration pass (igm_energy_pass2_f32) against two one-iteration
passes (igm_energy_pass_f32, mode 2), and the solver on pairs against the
solver on single iterations.

The two kernels run the same level sweep and Adam arithmetic on the same
float32 operands, so U, V and the moments must agree bit for bit.  The COSTs
are float64 sums of the same float32 cell energies grouped into different
blocks (60 against 62 columns per wave), so they agree to float64 rounding:
rel 1e-12.  The one-iteration pass itself is held to the oracle in
test_gpu_energy.py and test_gpu_solver.py.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from igm_amd import _lib as L
from igm_amd.energy import EnergyOperator
from igm_amd.solver import LegacyAdamState, adam_schedule, solve
from oracle import energy as E
from tests.helpers import dzeta_for, small_glacier, smooth_velocity, to_dev

pytestmark = pytest.mark.gpu
PHYS = E.DEFAULT_PHYSICS


def gpu_op(g, nz, dev):
    fields = {k: to_dev(g[k], dev) for k in ("thk", "usurf", "arrhenius", "slidingco", "dX")}
    return EnergyOperator(dzeta_for(nz), PHYS, fields=fields, dtype=torch.float32)


def es_device(dev):
    """A device EarlyStopping state that never stops (records the COSTs)."""
    es = L.IgmEarlyStop(best_loss=0.0, relative_min_delta=2e-4, has_best=0, wait=0,
                        patience=2**30, stopped=0, stop_iter=-1, iters_done=0)
    return torch.frombuffer(bytearray(bytes(es)), dtype=torch.uint8).to(dev)


def random_moments(nz, ny, nx, seed, dev):
    rng = np.random.default_rng(seed)
    m = np.empty((nz, ny, nx, 4))
    m[..., 0] = 0.1 * rng.standard_normal((nz, ny, nx))
    m[..., 1] = 0.01 * np.abs(rng.standard_normal((nz, ny, nx)))
    m[..., 2] = 0.1 * rng.standard_normal((nz, ny, nx))
    m[..., 3] = 0.01 * np.abs(rng.standard_normal((nz, ny, nx)))
    return to_dev(m, dev)


def first_mismatch(a, b):
    d = (a != b).nonzero()
    if d.numel() == 0:
        return None
    idx = tuple(int(x) for x in d[0])
    return idx, float(a[idx]), float(b[idx]), float((a.double() - b.double()).abs().max())


# (ny, nx, nz): ragged widths (not multiples of 60 or 62), few rows per
# wave, a single strip, tiny grids
CASES = [(100, 100, 10), (37, 131, 5), (257, 61, 2), (8, 300, 3), (5, 7, 4), (130, 181, 8),
         (64, 120, 6), (2, 2, 10), (3, 62, 10), (61, 60, 10), (50, 70, 12)]


@pytest.mark.parametrize("ny,nx,nz", CASES)
def test_pair_equals_two_single_passes(dev, ny, nx, nz):
    print(f"case pair vs singles ny={ny} nx={nx} nz={nz}", flush=True)
    g = small_glacier(ny, nx, seed=ny * 7 + nx)
    U, V = smooth_velocity(nz, ny, nx, seed=nz + ny)
    op = gpu_op(g, nz, dev)
    assert op.pair_supported
    U0, V0 = to_dev(U, dev), to_dev(V, dev)
    mom0 = random_moments(nz, ny, nx, ny + nx, dev)
    st = LegacyAdamState(lr=3.0)
    a1 = L.IgmAdam(alpha=st.alpha(5), beta1=0.9, beta2=0.999, epsilon=1e-7)
    st.lr = 2.0
    a2 = L.IgmAdam(alpha=st.alpha(6), beta1=0.9, beta2=0.999, epsilon=1e-7)
    hist_s = torch.zeros(2, dtype=torch.float64, device=dev)
    hist_p = torch.zeros(2, dtype=torch.float64, device=dev)

    # two one-iteration passes (moments in place)
    ms = mom0.clone()
    U1, V1, U2, V2 = (torch.empty_like(U0) for _ in range(4))
    es_s = es_device(dev)
    op.run_pass(2, U0, V0, Uo=U1, Vo=V1, moments=ms, adam=a1)
    op.finalize(es=es_s.data_ptr(), cost_hist=hist_s, it=0)
    op.run_pass(2, U1, V1, Uo=U2, Vo=V2, moments=ms, adam=a2)
    op.finalize(es=es_s.data_ptr(), cost_hist=hist_s, it=1)

    # one two-iteration pass
    mi, mo = mom0.clone(), torch.full_like(mom0, float("nan"))
    Up, Vp = torch.full_like(U0, float("nan")), torch.full_like(V0, float("nan"))
    part = torch.zeros_like(op.partials)
    es_p = es_device(dev)
    op.run_pass2(U0, V0, Up, Vp, mi, mo, a1, a2, part)
    op.finalize_pair(es=es_p.data_ptr(), cost_hist=hist_p, it=0, partials=part)
    torch.cuda.synchronize()

    assert torch.equal(mi, mom0), "the pass must not write its input moments"
    for name, a, b in (("U", Up, U2), ("V", Vp, V2), ("moments", mo, ms)):
        assert first_mismatch(a, b) is None, (name, first_mismatch(a, b))
    hs, hp = hist_s.cpu().numpy(), hist_p.cpu().numpy()
    assert np.all(np.abs(hp - hs) <= 1e-12 * np.abs(hs)), (hp, hs)


@pytest.mark.parametrize("ny,nx,nz", [(100, 100, 10), (37, 131, 5), (5, 7, 4)])
def test_pair_without_input_moments_is_a_fresh_optimizer(dev, ny, nx, nz):
    """moments_in = NULL stands for zero moments: the same U, V, moments and
    COSTs, bit for bit, as a pass reading a zeroed buffer; moments_out = NULL
    drops the moments and keeps U, V and the COSTs."""
    g = small_glacier(ny, nx, seed=ny + 2 * nx)
    U, V = smooth_velocity(nz, ny, nx, seed=nz + 3 * ny)
    op = gpu_op(g, nz, dev)
    U0, V0 = to_dev(U, dev), to_dev(V, dev)
    st = LegacyAdamState(lr=3.0)
    a1 = L.IgmAdam(alpha=st.alpha(1), beta1=0.9, beta2=0.999, epsilon=1e-7)
    a2 = L.IgmAdam(alpha=st.alpha(2), beta1=0.9, beta2=0.999, epsilon=1e-7)
    outs = []
    for zero_in in (torch.zeros((nz, ny, nx, 4), device=dev), None):
        mo = torch.full((nz, ny, nx, 4), float("nan"), device=dev)
        Uo, Vo = torch.full_like(U0, float("nan")), torch.full_like(V0, float("nan"))
        part = torch.zeros_like(op.partials)
        hist = torch.zeros(2, dtype=torch.float64, device=dev)
        es = es_device(dev)
        op.run_pass2(U0, V0, Uo, Vo, zero_in, mo, a1, a2, part)
        op.finalize_pair(es=es.data_ptr(), cost_hist=hist, it=0, partials=part)
        outs.append((Uo, Vo, mo, hist))
    # moments_out = NULL: the moments are dropped; U, V and the COSTs are kept
    Uo, Vo = torch.full_like(U0, float("nan")), torch.full_like(V0, float("nan"))
    part = torch.zeros_like(op.partials)
    hist = torch.zeros(2, dtype=torch.float64, device=dev)
    es = es_device(dev)
    op.run_pass2(U0, V0, Uo, Vo, None, None, a1, a2, part)
    op.finalize_pair(es=es.data_ptr(), cost_hist=hist, it=0, partials=part)
    torch.cuda.synchronize()
    for name, a, b in zip(("U", "V", "moments", "COST"), outs[0], outs[1]):
        assert first_mismatch(a, b) is None, (name, first_mismatch(a, b))
    for name, a, b in (("U", Uo, outs[0][0]), ("V", Vo, outs[0][1]), ("COST", hist, outs[0][3])):
        assert first_mismatch(a, b) is None, (name, first_mismatch(a, b))


@pytest.mark.parametrize("ny,nx,nz", [(100, 100, 10), (37, 131, 5), (5, 7, 4), (61, 60, 12)])
def test_pair_interleaved_layouts(dev, ny, nx, nz):
    """V (Vout) = NULL: U (Uout) holds (U, V) interleaved per node and level.
    All four input/output layouts give the planar pass's U, V, moments and
    COSTs bit for bit."""
    g = small_glacier(ny, nx, seed=3 * ny + nx)
    U, V = smooth_velocity(nz, ny, nx, seed=nz + 5 * ny)
    op = gpu_op(g, nz, dev)
    U0, V0 = to_dev(U, dev), to_dev(V, dev)
    UV0 = torch.stack([U0, V0], dim=-1).contiguous()
    mom0 = random_moments(nz, ny, nx, ny * nx, dev)
    st = LegacyAdamState(lr=3.0)
    a1 = L.IgmAdam(alpha=st.alpha(3), beta1=0.9, beta2=0.999, epsilon=1e-7)
    a2 = L.IgmAdam(alpha=st.alpha(4), beta1=0.9, beta2=0.999, epsilon=1e-7)

    def run(in_il, out_il):
        mo = torch.full_like(mom0, float("nan"))
        if out_il:
            UVo = torch.full((nz, ny, nx, 2), float("nan"), device=dev)
            dst = (UVo, None)
        else:
            dst = (torch.full_like(U0, float("nan")), torch.full_like(V0, float("nan")))
        part = torch.zeros_like(op.partials)
        hist = torch.zeros(2, dtype=torch.float64, device=dev)
        es = es_device(dev)
        op.run_pass2(*((UV0, None) if in_il else (U0, V0)), *dst, mom0, mo, a1, a2, part)
        op.finalize_pair(es=es.data_ptr(), cost_hist=hist, it=0, partials=part)
        torch.cuda.synchronize()
        Uo, Vo = (dst[0][..., 0], dst[0][..., 1]) if out_il else dst
        return Uo, Vo, mo, hist

    ref = run(False, False)
    for in_il, out_il in ((True, False), (False, True), (True, True)):
        got = run(in_il, out_il)
        for name, a, b in zip(("U", "V", "moments", "COST"), got, ref):
            assert first_mismatch(a, b) is None, (in_il, out_il, name, first_mismatch(a, b))


def test_pair_errors(dev):
    g = small_glacier(20, 30, seed=1)
    op = gpu_op(g, 10, dev)
    U0 = torch.zeros((10, 20, 30), device=dev)
    m = torch.zeros((10, 20, 30, 4), device=dev)
    a = L.IgmAdam(alpha=1.0, beta1=0.9, beta2=0.999, epsilon=1e-7)
    with pytest.raises(L.IgmError, match="another"):
        op.run_pass2(U0, U0.clone(), torch.empty_like(U0), torch.empty_like(U0), m, m, a, a,
                     op.partials)
    b = L.IgmAdam(alpha=1.0, beta1=0.8, beta2=0.999, epsilon=1e-7)
    with pytest.raises(L.IgmError, match="share"):
        op.run_pass2(U0, U0.clone(), torch.empty_like(U0), torch.empty_like(U0), m, m.clone(), a,
                     b, op.partials)
    op16 = gpu_op(g, 16, dev)
    assert not op16.pair_supported


@pytest.mark.parametrize("nit", [1, 2, 7, 20, 33])
def test_solver_pairs_equal_single_iterations(dev, nit):
    """solve() on pairs (odd counts start with a one-iteration pass) gives
    the single-iteration solve's field, COST history and step counter."""
    print(f"case solver pairs nit={nit}", flush=True)
    g = small_glacier(90, 140, seed=3)
    op = gpu_op(g, 10, dev)
    z = torch.zeros((10, 90, 140), device=dev)
    sched = adam_schedule(30.0, nit // 3, nit - nit // 3)
    o1, o2 = LegacyAdamState(lr=1.0), LegacyAdamState(lr=1.0)
    o1.iterations = o2.iterations = 4  # a persistent step counter, as across update() calls
    r1 = solve(op, z, z.clone(), nit, o1, early_stopping=False, lr_schedule=sched, pairs=False)
    r2 = solve(op, z, z.clone(), nit, o2, early_stopping=False, lr_schedule=sched, pairs=True)
    assert r1.iterations == r2.iterations == nit and o1.iterations == o2.iterations == 4 + nit
    assert first_mismatch(r2.U, r1.U) is None and first_mismatch(r2.V, r1.V) is None
    c1, c2 = np.asarray(r1.costs), np.asarray(r2.costs)
    assert c1.shape == c2.shape == (nit,)
    assert np.all(np.abs(c2 - c1) <= 1e-12 * np.abs(c1)), (c1, c2)


@pytest.mark.parametrize("patience,delta,nit,warm", [(1, 2e-2, 60, 0), (2, 5e-3, 80, 0),
                                                      (3, 2e-3, 100, 0), (10, 2e-4, 100, 0),
                                                      (10, 2e-4, 7, 0), (3, 2e-2, 100, 0),
                                                      (10, 5e-2, 100, 0), (10, 2e-4, 100, 80),
                                                      (4, 1e-3, 100, 80), (3, 5e-4, 100, 80),
                                                      (5, 3e-4, 100, 80), (2, 3e-4, 100, 40)])
def test_solver_with_early_stopping_on_pairs(dev, patience, delta, nit, warm):
    """With early stopping the device decides, pass by pass, whether two
    iterations are safe (the first cannot end the solve) or one: the result
    is the one-iteration solver's -- same stop, field, COSTs and counter."""
    print(f"case solver pairs + early stopping patience={patience} delta={delta} nit={nit} warm={warm}",
          flush=True)
    g = small_glacier(90, 140, seed=5)
    op = gpu_op(g, 10, dev)
    z = torch.zeros((10, 90, 140), device=dev)
    U0, V0 = z, z.clone()
    if warm:  # a warm start, as in a transient run's update()
        r0 = solve(op, z, z.clone(), warm, LegacyAdamState(lr=1.0), early_stopping=False, pairs=False)
        U0, V0 = r0.U, r0.V
    o1, o2 = LegacyAdamState(lr=1.0), LegacyAdamState(lr=1.0)
    kw = dict(early_stopping=True, relative_min_delta=delta, patience=patience, check_every=4)
    r1 = solve(op, U0, V0, nit, o1, pairs=False, **kw)
    r2 = solve(op, U0, V0, nit, o2, pairs=True, **kw)
    print(f"  single: {r1.iterations} its stopped {r1.stopped}; pairs: {r2.iterations} its stopped {r2.stopped}")
    assert (r2.iterations, r2.stopped) == (r1.iterations, r1.stopped)
    assert o1.iterations == o2.iterations == r1.iterations
    assert first_mismatch(r2.U, r1.U) is None and first_mismatch(r2.V, r1.V) is None
    c1, c2 = np.asarray(r1.costs), np.asarray(r2.costs)
    assert c1.shape == c2.shape
    assert np.all(np.abs(c2 - c1) <= 1e-12 * np.abs(c1)), (c1, c2)
