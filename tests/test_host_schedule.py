"""The solver's step schedule (solver.adam_schedule): geometric warm-up,
flat part and geometric decay, each checked against its closed form; with no
warm-up it is the flat-then-decay schedule, value for value."""
import math

from igm_amd.solver import adam_schedule


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
