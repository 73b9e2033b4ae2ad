"""The solver's step schedule (solver.adam_schedule): geometric warm-up,
flat part and geometric decay, each checked against its closed form; with no
warm-up it is the flat-then-decay schedule, value for value.  The Adam step
tables built from it (solver.step_table): TF legacy Adam's bias-corrected
step in float32, kept per schedule key and reused."""
import math

import numpy as np
import torch

from igm_amd import solver as slv
from igm_amd.solver import LegacyAdamState, adam_schedule, step_table


def test_flat_then_decay_closed_form():
    f = adam_schedule(150.0, 40, 60, 30.0)
    assert all(f(i) == 150.0 for i in range(40))
    for i in range(40, 100):
        assert math.isclose(f(i), 150.0 * (30.0 / 150.0) ** ((i - 40) / 60), rel_tol=1e-12)
    assert math.isclose(f(100), 30.0, rel_tol=1e-12) and f(500) == 30.0  # held at the end step


def test_warmup_ramp_flat_decay():
    n_w, n_f, n_d = 17, 35, 122
    f = adam_schedule(200.0, n_f, n_d, 30.0, n_warmup=n_w, lr_start=50.0)
    for i in range(n_w):  # geometric from 50 towards 200
        assert math.isclose(f(i), 50.0 * 4.0 ** (i / n_w), rel_tol=1e-12)
    assert all(f(i) == 200.0 for i in range(n_w, n_w + n_f))
    for i in range(n_w + n_f, n_w + n_f + n_d):
        assert math.isclose(f(i), 200.0 * (30.0 / 200.0) ** ((i - n_w - n_f) / n_d), rel_tol=1e-12)
    # monotone up, then down
    v = [f(i) for i in range(n_w + n_f + n_d)]
    assert all(a < b for a, b in zip(v[:n_w], v[1:n_w + 1]))
    assert all(a >= b for a, b in zip(v[n_w + n_f:], v[n_w + n_f + 1:]))


def test_no_warmup_is_the_plain_schedule():
    g = adam_schedule(120.0, 57, 133, 10.0)
    for kw in ({"n_warmup": 0}, {"n_warmup": 0, "lr_start": 3.0}):
        h = adam_schedule(120.0, 57, 133, 10.0, **kw)
        assert all(g(i) == h(i) for i in range(200))


def _alpha_closed_form(lr, b1, b2, t):
    # lr * sqrt(1 - beta2^t) / (1 - beta1^t) in float32 (equal to a few ulps:
    # float32 powers in another order)
    f = np.float32
    return f(lr) * np.sqrt(f(1) - f(b2) ** f(t)) / (f(1) - f(b1) ** f(t))


def test_step_table_is_the_bias_corrected_step():
    opt = LegacyAdamState(lr=7.0, beta1=0.93, beta2=0.95, epsilon=1e-9)
    sch = adam_schedule(100.0, 20, 40, 10.0, n_warmup=6, lr_start=100.0 / 6.0)
    for t0, s in ((0, sch), (37, sch), (5, None), (11, lambda i: 3.0 + i)):
        tab = step_table(opt, t0, 70, s, torch.float32)
        assert tab.dtype == np.float32 and tab.shape == (70,)
        for i in range(70):
            lr = 7.0 if s is None else s(i)
            assert math.isclose(float(tab[i]), float(_alpha_closed_form(lr, 0.93, 0.95, t0 + 1 + i)), rel_tol=2e-6)


def test_step_tables_are_kept_per_key():
    opt = LegacyAdamState(lr=1.0, beta1=0.93, beta2=0.95)
    mk = lambda: adam_schedule(100.0, 50, 110, 10.0, n_warmup=8, lr_start=100.0 / 6.0)
    a = step_table(opt, 0, 168, mk(), torch.float32)
    assert step_table(opt, 0, 168, mk(), torch.float32) is a  # an equal schedule: the same table
    assert not a.flags.writeable
    assert step_table(opt, 0, 166, mk(), torch.float32) is not a  # another length
    assert step_table(LegacyAdamState(lr=1.0, beta1=0.95, beta2=0.95), 0, 168, mk(), torch.float32) is not a
    f = lambda i: 100.0  # no key: built every time
    assert step_table(opt, 0, 10, f, torch.float32) is not step_table(opt, 0, 10, f, torch.float32)
    # a polish every time step, each followed by a check whose optimizer counter
    # grows: the polish's table stays (least recently used goes first)
    chk = LegacyAdamState(lr=1.0)
    for step in range(3 * slv._TABLES_MAX):
        assert step_table(opt, 0, 168, mk(), torch.float32) is a
        step_table(chk, 10 * step, 10, None, torch.float32)
    assert len(slv._TABLES) <= slv._TABLES_MAX
