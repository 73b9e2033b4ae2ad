"""Seeded synthetic glaciers: the input fields of the five workloads.

Every field is float32 numpy on the host, shaped (Ny, Nx), with x_i = i dx,
y_j = j dx.  Recipes:

* ``grf(shape, sigma, ell, dx, seed)``: a Gaussian random field --
  ``default_rng(seed).standard_normal`` filtered in Fourier space by a
  Gaussian kernel of standard deviation ell/dx cells (periodic), de-meaned
  and scaled to standard deviation sigma.
* a dome is ``Hpeak * max(0, 1 - r^2/R^2)``; several domes combine by max.
* ``usurf = bed + thk``; the bed is ``2500 - 0.1 x + grf(20 m, 500 m)``.

Workloads (sizes and physics of each):

* ``dome100``  -- 100 x 100 at dx=100 m, one dome R=4 km, Hpeak=300 m,
  centred; A=78, C=0.0464 uniform (the CPU-runnable plain solve).
* ``domes400`` -- 400 x 400 at dx=50 m, 4 domes R~U[2,6] km,
  Hpeak~U[150,450] m; per-cell A~U[50,100]; C=0.0464.
* ``domes1000`` -- 1000 x 1000 at dx=50 m, 20 domes, same ranges; A=78,
  C=0.0464.
* ``patch_batch`` -- a training batch of square patches with sides
  ~U{128..512}, Hpeak up to 500 m, C log-uniform in [0.01, 0.1].
* ``domes1500`` -- 1500 x 1500 at dx=50 m for transient runs.
"""
import numpy as np

DEFAULT_A = 78.0
DEFAULT_C = 0.0464


def grf(shape, sigma, ell, dx, seed):
    ny, nx = shape
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((ny, nx))
    s = ell / dx  # kernel std in cells
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.fftfreq(nx)[None, :]
    filt = np.exp(-2.0 * (np.pi ** 2) * (s ** 2) * (kx ** 2 + ky ** 2))
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    f -= f.mean()
    std = f.std()
    if std > 0:
        f *= sigma / std
    return f


def domes(shape, dx, centres, radii, peaks):
    ny, nx = shape
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    X, Y = np.meshgrid(x, y)
    H = np.zeros(shape)
    for (cx, cy), R, Hp in zip(centres, radii, peaks):
        r2 = (X - cx) ** 2 + (Y - cy) ** 2
        H = np.maximum(H, Hp * np.maximum(0.0, 1.0 - r2 / R ** 2))
    return H


def _bed(shape, dx, seed, slope=0.1, base=2500.0):
    ny, nx = shape
    x = np.arange(nx) * dx
    return base - slope * x[None, :] + grf(shape, 20.0, 500.0, dx, seed)


def _pack(bed, thk, arrhenius, slidingco, dx):
    f = np.float32
    thk = thk.astype(f)
    bed = bed.astype(f)
    shape = thk.shape
    A = np.broadcast_to(np.asarray(arrhenius, f), shape).astype(f)
    C = np.broadcast_to(np.asarray(slidingco, f), shape).astype(f)
    return dict(thk=thk, usurf=(bed + thk).astype(f), topg=bed, arrhenius=A, slidingco=C,
                dX=np.full(shape, dx, f), dx=float(dx))


def random_domes(shape, dx, n, seed, rmin=2000.0, rmax=6000.0, hmin=150.0, hmax=450.0, margin=0.15):
    ny, nx = shape
    rng = np.random.default_rng(seed)
    cx = rng.uniform(margin, 1 - margin, n) * (nx - 1) * dx
    cy = rng.uniform(margin, 1 - margin, n) * (ny - 1) * dx
    R = rng.uniform(rmin, rmax, n)
    Hp = rng.uniform(hmin, hmax, n)
    return domes(shape, dx, list(zip(cx, cy)), R, Hp)


def dome100():
    shape, dx = (100, 100), 100.0
    bed = _bed(shape, dx, seed=0)
    c = ((shape[1] - 1) * dx / 2, (shape[0] - 1) * dx / 2)
    H = domes(shape, dx, [c], [4000.0], [300.0])
    return _pack(bed, H, DEFAULT_A, DEFAULT_C, dx)


def domes400(seed=1):
    shape, dx = (400, 400), 50.0
    bed = _bed(shape, dx, seed=seed)
    H = random_domes(shape, dx, 4, seed=seed)
    A = np.random.default_rng(100 + seed).uniform(50.0, 100.0, shape)
    return _pack(bed, H, A, DEFAULT_C, dx)


def domes1000(seed=2):
    shape, dx = (1000, 1000), 50.0
    bed = _bed(shape, dx, seed=seed)
    H = random_domes(shape, dx, 20, seed=seed)
    return _pack(bed, H, DEFAULT_A, DEFAULT_C, dx)


def domes1500(seed=4):
    shape, dx = (1500, 1500), 50.0
    bed = _bed(shape, dx, seed=seed)
    H = random_domes(shape, dx, 40, seed=seed)
    return _pack(bed, H, DEFAULT_A, DEFAULT_C, dx)


def glacier_patch(side, seed, dx=50.0, hmax=500.0):
    """One training patch: random bed slope/roughness, 1-6 domes, C log-uniform
    in [0.01, 0.1], A = 78 or per-cell U[50, 100] (half of the patches each)."""
    rng = np.random.default_rng(seed)
    shape = (side, side)
    slope = rng.uniform(0.0, 0.2)
    ang = rng.uniform(0, 2 * np.pi)
    x = np.arange(side) * dx
    X, Y = np.meshgrid(x, x)
    bed = 2500.0 - slope * (np.cos(ang) * X + np.sin(ang) * Y) + grf(shape, rng.uniform(5, 40), rng.uniform(200, 1000), dx, seed + 7919)
    ext = (side - 1) * dx
    n = int(rng.integers(1, 7))
    rmax = max(min(6000.0, 0.6 * ext), 800.0)
    H = random_domes(shape, dx, n, seed=seed + 104729, rmin=min(1000.0, rmax * 0.5), rmax=rmax,
                     hmin=50.0, hmax=hmax, margin=0.1)
    C = float(np.exp(rng.uniform(np.log(0.01), np.log(0.1))))
    if rng.uniform() < 0.5:
        A = DEFAULT_A
    else:
        A = rng.uniform(50.0, 100.0, shape)
    return _pack(bed, H, A, C, dx)


def patch_batch(step, n=32, smin=128, smax=512, dx=50.0):
    """A batch of ``n`` patches of uneven sizes, seeded by ``3 + step``."""
    rng = np.random.default_rng(3 + step)
    sides = rng.integers(smin, smax + 1, n)
    seeds = rng.integers(0, 2**31 - 1, n)
    return [glacier_patch(int(s), int(sd), dx=dx) for s, sd in zip(sides, seeds)]


WORKLOADS = {
    "dome100": dome100,
    "domes400": domes400,
    "domes1000": domes1000,
    "domes1500": domes1500,
}
