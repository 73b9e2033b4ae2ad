"""GPU transport kernel vs the oracle restatement of
compute_divflux_slope_limiter + thk.update (float32 both sides: the same op
order, so agreement to a few ulp), plus CFL and SMB."""
import numpy as np
import pytest
import torch

from igm_amd import _lib as L
from igm_amd import config
from igm_amd import synthetic as syn
from igm_amd.processes import smb_simple, thk as thkmod, time as timemod
from oracle import transport as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("slope,ny,nx", [("superbee", 60, 90), ("minmod", 33, 17), ("godunov", 5, 300),
                                         ("superbee", 1, 7),
                                         # several 60-column strips and 8-row blocks, with remainders
                                         ("superbee", 131, 245), ("minmod", 9, 61), ("godunov", 17, 121),
                                         ("superbee", 64, 1), ("superbee", 8, 120)])
def test_transport_matches_oracle(dev, slope, ny, nx):
    rng = np.random.default_rng(ny * nx)
    h = np.maximum(0, rng.normal(100, 120, (ny, nx))).astype(np.float32)
    u = rng.normal(0, 80, (ny, nx)).astype(np.float32)
    v = rng.normal(0, 80, (ny, nx)).astype(np.float32)
    smb = rng.normal(0, 2, (ny, nx)).astype(np.float32)
    topg = rng.normal(2400, 50, (ny, nx)).astype(np.float32)
    dx, dt = 50.0, 0.2
    t_ref, ls_ref, us_ref, div_ref = T.thk_update(h, u, v, smb, topg, dx, dt, slope)
    d = {k: torch.as_tensor(a, device=dev) for k, a in dict(h=h, u=u, v=v, smb=smb, topg=topg).items()}
    out = {k: torch.empty((ny, nx), device=dev) for k in ("thk", "ls", "us", "div")}
    L.call("igm_transport_f32", ny, nx, L.ptr(d["u"]), L.ptr(d["v"]), L.ptr(d["h"]), L.ptr(d["smb"]),
           L.ptr(d["topg"]), dx, dx, dt, {"godunov": 0, "minmod": 1, "superbee": 2}[slope], 0.91, 0.0,
           L.ptr(out["thk"]), L.ptr(out["ls"]), L.ptr(out["us"]), L.ptr(out["div"]), L.stream_ptr())
    np.testing.assert_allclose(out["div"].cpu().numpy(), div_ref, rtol=1e-5, atol=1e-5 * np.abs(div_ref).max())
    np.testing.assert_allclose(out["thk"].cpu().numpy(), t_ref, rtol=1e-5, atol=1e-3)
    np.testing.assert_allclose(out["us"].cpu().numpy(), us_ref, rtol=1e-6, atol=1e-3)


def test_time_and_smb_processes(dev):
    class S:
        pass
    cfg = config.default_cfg()
    cfg.processes.smb_simple.array = [["time", "gradabl", "gradacc", "ela", "accmax"],
                                      [0, 0.005, 0.005, 2800, 2.0], [3000, 0.005, 0.005, 2800, 2.0]]
    st = S()
    g = syn.dome100()
    st.thk = torch.as_tensor(g["thk"], device=dev)
    st.topg = torch.as_tensor(g["topg"], device=dev)
    st.dx = 100.0
    st.it = 0
    thkmod.initialize(cfg, st)
    timemod.initialize(cfg, st)
    smb_simple.initialize(cfg, st)
    st.ubar = torch.full_like(st.thk, 10.0); st.ubar[3, 4] = -300.0
    st.vbar = torch.zeros_like(st.thk)
    smb_simple.update(cfg, st)
    ref = T.smb_simple(st.usurf.cpu().numpy(), 0.005, 0.005, 2800.0, 2.0)
    np.testing.assert_allclose(st.smb.cpu().numpy(), ref, rtol=1e-6, atol=1e-6)
    timemod.update(cfg, st)
    # time.py:54-59: the first step snaps to the first save time (t = start), dt = 0
    assert st.dt == 0.0 and st.saveresult
    assert st.dt_target == pytest.approx(T.cfl_dt(st.ubar.cpu().numpy(), st.vbar.cpu().numpy(), 100.0, 0.3, 1.0),
                                         rel=1e-6)
    timemod.update(cfg, st)
    assert st.dt == pytest.approx(st.dt_target) and st.t == pytest.approx(2000.0 + st.dt_target)
    vol0 = float(st.thk.double().sum())
    thkmod.update(cfg, st)
    assert torch.isfinite(st.thk).all() and float(st.thk.min()) >= 0.0
    assert torch.equal(st.usurf, st.lsurf + st.thk)
