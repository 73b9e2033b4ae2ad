"""Retraining host logic of the emulator (CPU restatement, numpy).

Restates ``igm/processes/iceflow/emulate/emulate.py:283-308``
(``split_into_patches``, "sequential") and the per-patch learning rate of
``emulate.py:209-210``.  Pinned in ``tests/test_host_retrain.py`` by closed
forms: the number and size of the patches for the grid sizes the workloads
use, the rows and columns the integer division drops, the traversal order,
and the decay lr * lr_decay^(i/1000) at i = 0, 1000, 2000.
"""
import numpy as np


def split_into_patches(X, nbmax):
    """emulate.py:283-308 with split_patch_method "sequential": X (1, ny, nx, C)
    -> (P, 1, ly, lx, C), columns of patches outermost, rows inner; the
    remainder rows/columns of ``int(ny / sy)`` are dropped."""
    ny, nx = X.shape[1], X.shape[2]
    sy = ny // nbmax + 1
    sx = nx // nbmax + 1
    ly = int(ny / sy)
    lx = int(nx / sx)
    XX = []
    for i in range(sx):
        for j in range(sy):
            XX.append(X[:, j * ly:(j + 1) * ly, i * lx:(i + 1) * lx, :])
    return np.stack(XX, axis=0)


def retrain_lr(lr, lr_decay, i):
    """emulate.py:209-210: the learning rate of patch i of an epoch."""
    if lr_decay < 1:
        return lr * (lr_decay ** (i / 1000))
    return lr
