"""Gradient-descent solver of the energy: legacy Keras Adam + EarlyStopping.

Restates ``igm/processes/iceflow/solve/solve.py:22-119`` and
``igm/processes/iceflow/utils.py:80-111`` (EarlyStopping) on the CPU.

Adam follows TensorFlow's legacy ``ResourceApplyAdam`` (what
``tf.keras.optimizers.legacy.Adam`` calls, solve.py:18-20):
  alpha = lr * sqrt(1 - beta2^t) / (1 - beta1^t)     (t = optimizer step + 1)
  m += (g - m) * (1 - beta1);  v += (g*g - v) * (1 - beta2)
  x -= (m * alpha) / (sqrt(v) + epsilon)
with beta1=0.9, beta2=0.999, epsilon=1e-7 (Keras defaults).  The step
counter is the optimizer's and persists across solves (the optimizer is
created once in solve.py:11-20); moments are fresh per solve because the
variables are new tf.Variables (solve.py:27-28).

Pins: Adam is checked against a hand-computed 2-step sequence and against
torch.optim.Adam after the epsilon/bias-correction mapping; the solver is
checked to reach the analytic inclined-slab velocity and the float64
minimiser (gradient norm ~ 0).  Bitwise parity with TensorFlow's Eigen
reductions is not reachable without TensorFlow ("parity unpinned" for the
exact float32 rounding of the 500-step trajectory).
"""
import numpy as np
import torch

from . import energy as E

BETA1, BETA2, EPSILON = 0.9, 0.999, 1e-7


class LegacyAdam:
    """tf.keras.optimizers.legacy.Adam semantics (see module docstring)."""

    def __init__(self, lr=1.0, beta1=BETA1, beta2=BETA2, epsilon=EPSILON, iterations=0):
        self.lr, self.beta1, self.beta2, self.epsilon = lr, beta1, beta2, epsilon
        self.iterations = iterations

    def alpha(self, dtype):
        t = self.iterations + 1
        if dtype == torch.float32:
            f = np.float32
            b1p = f(np.power(f(self.beta1), f(t)))
            b2p = f(np.power(f(self.beta2), f(t)))
            return float(f(self.lr) * np.sqrt(f(1) - b2p) / (f(1) - b1p))
        b1p = self.beta1 ** t
        b2p = self.beta2 ** t
        return self.lr * np.sqrt(1 - b2p) / (1 - b1p)

    def apply(self, xs, gs, ms, vs):
        """One step on lists of tensors, in place."""
        a = self.alpha(xs[0].dtype)
        for x, g, m, v in zip(xs, gs, ms, vs):
            m += (g - m) * (1 - self.beta1)
            v += (g * g - v) * (1 - self.beta2)
            x -= (m * a) / (torch.sqrt(v) + self.epsilon)
        self.iterations += 1


class EarlyStopping:
    """iceflow/utils.py:80-111."""

    def __init__(self, relative_min_delta=1e-3, patience=10):
        self.relative_min_delta = relative_min_delta
        self.patience = patience
        self.wait = 0
        self.best_loss = None

    def should_stop(self, current_loss):
        if self.best_loss is None:
            self.best_loss = current_loss
            return False
        relative_improvement = (self.best_loss - current_loss) / abs(self.best_loss)
        if relative_improvement > self.relative_min_delta:
            self.best_loss = current_loss
            self.wait = 0
            return False
        self.wait += 1
        return self.wait >= self.patience


def solve_iceflow(U, V, fieldin, dzeta, thk, nbitmax=100, lr=1.0, optimizer=None,
                  early_stopping=True, phys=None):
    """solve.py:22-119: minimise the energy over (U, V) from the given start.

    U, V: (Nz, Ny, Nx) tensors (warm start); fieldin: list of (1, Ny, Nx).
    Returns (U, V masked by thk>0, list of COST values, iterations run).
    """
    U = U.detach().clone()
    V = V.detach().clone()
    mU, vU = torch.zeros_like(U), torch.zeros_like(U)
    mV, vV = torch.zeros_like(V), torch.zeros_like(V)
    opt = optimizer if optimizer is not None else LegacyAdam(lr=lr)
    es = EarlyStopping(relative_min_delta=0.0002, patience=10)
    costs = []
    it = 0
    for i in range(nbitmax):
        cost, _, gU, gV = E.energy_and_grad(U, V, fieldin, dzeta, phys)
        costs.append(float(cost))
        opt.apply([U, V], [gU, gV], [mU, mV], [vU, vV])
        it = i + 1
        if early_stopping and es.should_stop(float(cost)):
            break
    mask = torch.as_tensor(thk) > 0
    U = torch.where(mask, U, torch.zeros_like(U))
    V = torch.where(mask, V, torch.zeros_like(V))
    return U, V, costs, it
