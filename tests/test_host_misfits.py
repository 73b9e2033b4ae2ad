"""The reference's accuracy measures (igm_amd.fields.misfits) on the CPU.

Pinned by closed forms -- identical fields give 0; U, V scaled by (1 + e)
give relative L1 and L2 of exactly e; a constant difference d in every
level gives velbar L1 = RMS = |d| (the vertical weights sum to 1) -- and
checked against the oracle's restatement of pretraining.py:502-535 and
diagnostic.py:74-84 on random fields.
"""
import numpy as np
import pytest
import torch

from igm_amd.fields import misfits
from oracle import fields as OF
from oracle import vert_disc as VD


def fields(nz=10, ny=23, nx=31, seed=0):
    rng = np.random.default_rng(seed)
    thk = torch.as_tensor(np.maximum(0.0, rng.uniform(-50, 400, (ny, nx))))
    U = torch.as_tensor(rng.standard_normal((nz, ny, nx)) * 50)
    V = torch.as_tensor(rng.standard_normal((nz, ny, nx)) * 50)
    return thk, U, V


def test_closed_forms():
    nz, vs = 10, 4.0
    vw = VD.define_vertical_weight(nz, vs)
    thk, U, V = fields(nz)
    z = misfits(thk, U, V, U, V, vw, vs)
    assert z == {"rel_l1": 0.0, "rel_l2": 0.0, "velbar_l1": 0.0, "velbar_rms": 0.0}
    e = 0.037
    m = misfits(thk, U, V, (1 + e) * U, (1 + e) * V, vw, vs)
    assert abs(m["rel_l1"] - e) <= 1e-12 and abs(m["rel_l2"] - e) <= 1e-12
    d = 2.5
    m = misfits(thk, U, V, U - d, V + d, vw, vs)
    assert abs(m["velbar_l1"] - d) <= 1e-12 and abs(m["velbar_rms"] - d) <= 1e-12


@pytest.mark.parametrize("seed", [0, 1])
def test_against_the_oracle(seed):
    nz, vs = 10, 4.0
    vw = VD.define_vertical_weight(nz, vs)
    thk, U, V = fields(nz, seed=seed)
    _, Up, Vp = fields(nz, seed=seed + 10)
    Up, Vp = U + 0.1 * Up, V + 0.1 * Vp
    m = misfits(thk, U, V, Up, Vp, vw, vs)
    r1, r2, b1, b2 = OF.computemisfitall(thk, U, V, Up, Vp, vw, vs)
    for a, b in ((m["rel_l1"], r1), (m["rel_l2"], r2), (m["velbar_l1"], b1), (m["velbar_rms"], b2)):
        assert abs(a - b) <= 1e-12 * abs(b), (a, b)
