"""Mass transport, CFL time step and simple SMB (CPU restatement, numpy float32).

Restates ``igm/utils/gradient/compute_divflux_slope_limiter.py:3-65``,
``igm/processes/thk/thk.py:27-54``, ``igm/processes/time/time.py:28-65`` and
``igm/processes/smb_simple/smb_simple.py:24-50``.

Pins (``tests/test_oracle_transport.py``): exact conservation of the flux
divergence (it sums to the boundary fluxes, which vanish for ice away from
the edges), a uniform-velocity translation of a smooth profile against the
exact shifted profile, godunov = first-order upwind closed form, limiter
identities (minmod/superbee on monotone and extremal data), the order of
accuracy on a smooth front (2 with superbee, 1 with godunov) and the TVD
property of the limited schemes on square waves.
"""
import numpy as np

f32 = np.float32


def minmod(a, b):
    """compute_divflux_slope_limiter.py:3-4."""
    return np.where((np.abs(a) < np.abs(b)) & (a * b > 0.0), a,
                    np.where((np.abs(a) > np.abs(b)) & (a * b > 0.0), b, f32(0)))


def maxmod(a, b):
    """compute_divflux_slope_limiter.py:6-7."""
    return np.where((np.abs(a) < np.abs(b)) & (a * b > 0.0), b,
                    np.where((np.abs(a) > np.abs(b)) & (a * b > 0.0), a, f32(0)))


def compute_divflux_slope_limiter(u, v, h, dx, dy, dt, slope_type="superbee"):
    """compute_divflux_slope_limiter.py:9-65: upwind div(u h) with a limiter."""
    u = np.asarray(u, f32); v = np.asarray(v, f32); h = np.asarray(h, f32)
    dx, dy, dt = f32(dx), f32(dy), f32(dt)
    u = np.concatenate([u[:, 0:1], f32(0.5) * (u[:, :-1] + u[:, 1:]), u[:, -1:]], 1)
    v = np.concatenate([v[0:1, :], f32(0.5) * (v[:-1, :] + v[1:, :]), v[-1:, :]], 0)
    Hx = np.pad(h, [[0, 0], [2, 2]])
    Hy = np.pad(h, [[2, 2], [0, 0]])
    sigpx = (Hx[:, 2:] - Hx[:, 1:-1]) / dx
    sigmx = (Hx[:, 1:-1] - Hx[:, :-2]) / dx
    sigpy = (Hy[2:, :] - Hy[1:-1, :]) / dy
    sigmy = (Hy[1:-1, :] - Hy[:-2, :]) / dy
    if slope_type == "godunov":
        slopex = np.zeros_like(sigpx)
        slopey = np.zeros_like(sigpy)
    elif slope_type == "minmod":
        slopex = minmod(sigmx, sigpx)
        slopey = minmod(sigmy, sigpy)
    elif slope_type == "superbee":
        slopex = maxmod(minmod(sigpx, f32(2.0) * sigmx), minmod(sigmx, f32(2.0) * sigpx))
        slopey = maxmod(minmod(sigpy, f32(2.0) * sigmy), minmod(sigmy, f32(2.0) * sigpy))
    else:
        raise ValueError(slope_type)
    half, one = f32(0.5), f32(1.0)
    w = Hx[:, 1:-2] + half * dx * (one - u * dt / dx) * slopex[:, :-1]
    e = Hx[:, 2:-1] - half * dx * (one + u * dt / dx) * slopex[:, 1:]
    s = Hy[1:-2, :] + half * dy * (one - v * dt / dy) * slopey[:-1, :]
    n = Hy[2:-1, :] - half * dy * (one + v * dt / dy) * slopey[1:, :]
    Qx = u * np.where(u > 0, w, e)
    Qy = v * np.where(v > 0, s, n)
    return ((Qx[:, 1:] - Qx[:, :-1]) / dx + (Qy[1:, :] - Qy[:-1, :]) / dy).astype(f32)


def thk_update(thk, ubar, vbar, smb, topg, dx, dt, slope_type="superbee",
               ratio_density=0.910, sealevel=0.0):
    """thk.py:27-54: thk <- max(thk + dt (smb - divflux), 0); lsurf; usurf."""
    divflux = compute_divflux_slope_limiter(ubar, vbar, thk, dx, dx, dt, slope_type)
    thk = np.maximum(np.asarray(thk, f32) + f32(dt) * (np.asarray(smb, f32) - divflux), f32(0))
    lsurf = np.maximum(np.asarray(topg, f32), -f32(ratio_density) * thk + f32(sealevel))
    usurf = lsurf + thk
    return thk.astype(f32), lsurf.astype(f32), usurf.astype(f32), divflux


def cfl_dt(ubar, vbar, dx, cfl=0.3, step_max=1.0):
    """time.py:31-49: dt_target = min(cfl dx / max|velbar|, step_max)."""
    velomax = max(float(np.max(np.abs(ubar))), float(np.max(np.abs(vbar))))
    if cfl > 0 and velomax > 0:
        return float(min(f32(cfl) * f32(dx) / f32(velomax), f32(step_max)))
    return float(step_max)


def smb_simple(usurf, gradabl, gradacc, ela, maxacc):
    """smb_simple.py:40-42: ELA-gradient SMB clipped to [-100, maxacc]."""
    smb = np.asarray(usurf, f32) - f32(ela)
    smb = smb * np.where(smb < 0, f32(gradabl), f32(gradacc))
    return np.clip(smb, f32(-100), f32(maxacc)).astype(f32)
