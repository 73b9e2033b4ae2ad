"""Host logic of the emulator's retraining (emulate.py:155-308) on the CPU:
the patch windows of split_into_patches, the per-patch learning-rate decay,
and the diagnostic's misfit (diagnostic.py:74-84), each against the oracle's
restatement and pinned by closed forms."""
import numpy as np
import pytest
import torch

from igm_amd import emulator as EM
from igm_amd import iceflow
from oracle import emulate as OE
from oracle import fields as OF
from oracle import vert_disc as VD


@pytest.mark.parametrize("ny,nx,nbmax,np_,ly,lx", [
    (1500, 1500, 750, 9, 500, 500),   # sy = 3: three 500-row bands
    (1000, 1000, 750, 4, 500, 500),
    (1001, 999, 750, 4, 500, 499),    # the 1001st row and the 999th column are dropped
    (100, 100, 750, 1, 100, 100),
    (750, 751, 750, 4, 375, 375),     # ny = nbmax already splits in two
    (37, 90, 20, 10, 18, 18),
])
def test_split_into_patches_closed_forms(ny, nx, nbmax, np_, ly, lx):
    X = np.arange(ny * nx * 2, dtype=np.float64).reshape(1, ny, nx, 2)
    P = OE.split_into_patches(X, nbmax)
    assert P.shape == (np_, 1, ly, lx, 2)
    sy, sx = ny // nbmax + 1, nx // nbmax + 1
    # columns of patches outermost: patch k covers rows (k % sy) and columns (k // sy)
    for k in range(np_):
        j, i = k % sy, k // sy
        assert P[k, 0, 0, 0, 0] == X[0, j * ly, i * lx, 0]
    # rows and columns beyond sy*ly, sx*lx are in no patch
    assert ny - sy * ly < sy and nx - sx * lx < sx
    covered = np.zeros((ny, nx), bool)
    for (j0, j1, i0, i1) in EM.split_into_patches(ny, nx, nbmax):
        assert not covered[j0:j1, i0:i1].any()
        covered[j0:j1, i0:i1] = True
    assert covered[:sy * ly, :sx * lx].all() and not covered[sy * ly:].any() and not covered[:, sx * lx:].any()


@pytest.mark.parametrize("ny,nx,nbmax", [(1500, 1500, 750), (1001, 999, 750), (37, 90, 20), (64, 700, 100)])
def test_split_into_patches_matches_the_oracle(ny, nx, nbmax):
    """The module's windows cut the same patches, in the same order."""
    X = np.random.default_rng(0).standard_normal((1, ny, nx, 5))
    P = OE.split_into_patches(X, nbmax)
    wins = EM.split_into_patches(ny, nx, nbmax)
    assert len(wins) == P.shape[0]
    for k, (j0, j1, i0, i1) in enumerate(wins):
        assert np.array_equal(X[0, j0:j1, i0:i1], P[k, 0])


def test_retrain_lr_decay():
    """lr * lr_decay^(i/1000): lr at patch 0, lr * decay at patch 1000,
    lr * decay^2 at 2000; constant when lr_decay >= 1."""
    lr, d = 2e-5, 0.95
    assert OE.retrain_lr(lr, d, 0) == lr
    assert OE.retrain_lr(lr, d, 1000) == pytest.approx(lr * d, rel=1e-15)
    assert OE.retrain_lr(lr, d, 2000) == pytest.approx(lr * d * d, rel=1e-15)
    assert OE.retrain_lr(lr, 1.0, 500) == lr
    for i in (0, 1, 7, 999, 1000, 4321):
        for dec in (0.5, 0.95, 1.0, 1.2):
            assert EM.retrain_lr(lr, dec, i) == OE.retrain_lr(lr, dec, i)


def test_diagnostic_misfit_against_the_oracle():
    """iceflow.computemisfit (masked thk > 1, mean over both components of
    the velbar difference) equals the oracle's computemisfit on random
    fields, and gives |d| twice for a constant difference d in every level
    (the vertical weights sum to one)."""
    nz, vs = 10, 4.0
    vw = VD.define_vertical_weight(nz, vs)
    rng = np.random.default_rng(4)
    ny, nx = 31, 47
    thk = torch.as_tensor(np.maximum(0.0, rng.uniform(-100, 400, (ny, nx))))
    dU = torch.as_tensor(rng.standard_normal((nz, ny, nx)) * 30)
    dV = torch.as_tensor(rng.standard_normal((nz, ny, nx)) * 30)
    w = torch.as_tensor(vw.reshape(nz, 1, 1), dtype=torch.float64)
    l1, l2 = iceflow.computemisfit(thk, (dU * w).sum(0), (dV * w).sum(0))
    r1, r2 = OF.computemisfit(thk, dU, dV, vw)
    assert l1 == pytest.approx(r1, rel=1e-12) and l2 == pytest.approx(r2, rel=1e-12)
    d = 3.25
    c1, c2 = iceflow.computemisfit(thk, torch.full((ny, nx), d, dtype=torch.float64),
                                   torch.full((ny, nx), -d, dtype=torch.float64))
    assert c1 == pytest.approx(d, rel=1e-14) and c2 == pytest.approx(d, rel=1e-14)
    assert iceflow.computemisfit(torch.zeros((ny, nx)), dU[0], dV[0]) == (0.0, 0.0)
