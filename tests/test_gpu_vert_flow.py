"""GPU vertical velocity (vert_flow v2 kinematic) vs the oracle, float32 both
sides; the process module on a glacier."""
import numpy as np
import pytest
import torch

from igm_amd import _lib as L
from igm_amd import config
from igm_amd.processes import vert_flow as VFP
from oracle import vert_disc as VD
from oracle import vert_flow as VF
from tests.helpers import small_glacier, smooth_velocity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nz,ny,nx", [(10, 33, 70), (10, 2, 2), (4, 5, 300), (10, 64, 257), (20, 17, 9)])
def test_w_kinematic_matches_oracle(dev, nz, ny, nx):
    print(f"case vert_flow nz={nz} ny={ny} nx={nx}", flush=True)
    g = small_glacier(ny, nx, dx=50.0, seed=nz + ny)
    U, V = smooth_velocity(nz, ny, nx, seed=ny, scale=80.0)
    levels = np.asarray(VD.compute_levels(nz, 4.0), np.float64)
    w = np.asarray(VD.define_vertical_weight(nz, 4.0), np.float64).ravel()
    topg = (g["usurf"] - g["thk"]).astype(np.float32)
    ref = VF.w_kinematic(U, V, topg, g["thk"], levels, w, 50.0)
    import ctypes as C
    lv = (C.c_double * nz)(*levels)
    vw = (C.c_double * nz)(*w)
    d = {k: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)
         for k, a in dict(U=U, V=V, topg=topg, thk=g["thk"]).items()}
    W = torch.empty((nz, ny, nx), device=dev)
    wb = torch.empty((ny, nx), device=dev)
    ws = torch.empty((ny, nx), device=dev)
    L.call("igm_vert_flow_kinematic_f32", nz, ny, nx, L.ptr(d["U"]), L.ptr(d["V"]), L.ptr(d["topg"]),
           L.ptr(d["thk"]), C.addressof(lv), C.addressof(vw), 50.0, L.ptr(W), L.ptr(wb), L.ptr(ws),
           L.stream_ptr())
    Wc = W.cpu().numpy()
    scale = float(np.abs(U).max()) * 0.2  # |U| x a typical slope: the size of the cancelling terms
    np.testing.assert_allclose(Wc, ref, rtol=1e-4, atol=2e-5 * scale)
    assert np.array_equal(wb.cpu().numpy(), Wc[0]) and np.array_equal(ws.cpu().numpy(), Wc[-1])


def test_process_module(dev):
    cfg = config.default_cfg()

    class S:
        pass
    st = S()
    nz, ny, nx = 10, 40, 60
    g = small_glacier(ny, nx, dx=100.0, seed=3)
    U, V = smooth_velocity(nz, ny, nx, seed=5)
    st.U = torch.as_tensor(U, dtype=torch.float32, device=dev)
    st.V = torch.as_tensor(V, dtype=torch.float32, device=dev)
    st.thk = torch.as_tensor(g["thk"], device=dev)
    st.topg = torch.as_tensor(g["usurf"] - g["thk"], device=dev)
    st.dx = 100.0
    VFP.initialize(cfg, st)
    VFP.update(cfg, st)
    levels = np.asarray(VD.compute_levels(nz, 4.0), np.float64)
    w = np.asarray(VD.define_vertical_weight(nz, 4.0), np.float64).ravel()
    ref = VF.w_kinematic(U, V, st.topg.cpu().numpy(), g["thk"], levels, w, 100.0)
    np.testing.assert_allclose(st.W.cpu().numpy(), ref, rtol=1e-4, atol=2e-5 * float(np.abs(U).max()) * 0.2)
    assert st.wvelbase.shape == (ny, nx) and st.wvelsurf.shape == (ny, nx)
    cfg.processes.vert_flow = type("C", (), {"version": 1, "method": "kinematic"})()
    with pytest.raises(NotImplementedError):
        VFP.update(cfg, st)
