"""Physics-informed emulator training on synthetic glaciers.

The analogue of igm/processes/pretraining/pretraining.py:198-493
(train_iceflow_emulator) with generated data instead of netCDF datasets:
each step draws a batch of glacier patches of uneven sizes, packs them into
one zero-padded NHWC batch with per-entry extents (the kernels mask
everything outside an entry's extent), and takes one legacy-Adam step on
the sum over the batch of the Blatter-Pattyn energy of the predicted field
(pretraining.py:318-338: loss = energy, no labels).  Augmentation: random
flips and transposition of each patch (pretraining.py:550-559).

CLI: ``python -m igm_amd.pretraining --steps 2000 --out <file>`` trains
from IGM's shipped weights (or from scratch with --scratch) and saves a
safetensors emulator.
"""
import argparse
import os
import time

import numpy as np
import torch

from . import emulator as EM
from . import synthetic as syn
from .energy import EnergyOperator
from .vert_disc import compute_levels, compute_zeta_dzeta

KEYS = ("thk", "usurf", "arrhenius", "slidingco", "dX")


def patch_tensor(p, dev):
    return torch.stack([torch.as_tensor(p[k]) for k in KEYS], -1).to(dev)


def make_pool(n, seed=3, smin=128, smax=512, dev=None, crops_from=(), crop_frac=0.5):
    """n patches: glacier_patch() draws, plus crops of large multi-dome fields
    (``crops_from``: list of field dicts) for a fraction ``crop_frac``."""
    dev = dev or torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(n):
        side = int(rng.integers(smin, smax + 1))
        if crops_from and rng.uniform() < crop_frac:
            src = crops_from[int(rng.integers(len(crops_from)))]
            H, W = src["thk"].shape
            side = min(side, H, W)
            j0 = int(rng.integers(0, H - side + 1))
            i0 = int(rng.integers(0, W - side + 1))
            p = {k: np.ascontiguousarray(src[k][j0:j0 + side, i0:i0 + side]) for k in KEYS}
        else:
            p = syn.glacier_patch(side, int(rng.integers(0, 2 ** 31 - 1)))
        pool.append(patch_tensor(p, dev))
    return pool


def augment(t, code):
    """code bits: 1 flip x, 2 flip y, 4 transpose."""
    if code & 1:
        t = torch.flip(t, dims=[1])
    if code & 2:
        t = torch.flip(t, dims=[0])
    if code & 4:
        t = t.transpose(0, 1)
    return t


class Batcher:
    def __init__(self, pool, batch, seed, dev):
        self.pool, self.batch, self.dev = pool, batch, dev
        self.rng = np.random.default_rng(seed)
        self.smax = max(max(p.shape[0], p.shape[1]) for p in pool)

    def next(self):
        idx = self.rng.integers(0, len(self.pool), self.batch)
        codes = self.rng.integers(0, 8, self.batch)
        parts = [augment(self.pool[i], int(c)) for i, c in zip(idx, codes)]
        H = max(p.shape[0] for p in parts)
        W = max(p.shape[1] for p in parts)
        X = torch.zeros((self.batch, H, W, 5), dtype=torch.float32, device=self.dev)
        ext = torch.empty((self.batch, 2), dtype=torch.int32, pin_memory=True)
        for b, p in enumerate(parts):
            X[b, :p.shape[0], :p.shape[1]] = p
            ext[b, 0], ext[b, 1] = p.shape[0], p.shape[1]
        # pinned host buffer + non-blocking copy: no host/device synchronisation per step
        return X, ext.to(self.dev, non_blocking=True)


def train(cnn, pool, steps, dzeta, phys, batch=32, lr=1e-4, lr_end=None, clipnorm=1.0, epsilon=1e-7,
          seed=3, log_every=100, val=None, log=print):
    """Energy-as-loss training; returns the list of (step, mean COST per entry).

    lr decays geometrically from ``lr`` to ``lr_end`` (constant if None).
    ``val(cnn)`` (optional) is called every log_every steps and its result logged."""
    dev = cnn.device
    bt = Batcher(pool, batch, seed, dev)
    hist = []
    t0 = time.time()
    for s in range(steps):
        X, ext = bt.next()
        op = EnergyOperator(dzeta, phys, X=X, extents=ext)
        fields = EM.nhwc_fields(X)
        Y, acts = cnn.forward(fields, extents=ext, keep=True)
        e, gY = op.energy_grad(Y)
        cnn.backward(gY, acts, fields, extents=ext)
        lr_s = lr if lr_end is None else lr * (lr_end / lr) ** (s / max(1, steps - 1))
        cnn.adam_step(lr_s, clipnorm=clipnorm, epsilon=epsilon)
        if (s + 1) % log_every == 0 or s == 0:
            c = float(e[:, 3].mean())
            hist.append((s + 1, c))
            msg = f"step {s + 1}: mean COST {c:.5f} lr {lr_s:.2e} ({time.time() - t0:.1f}s)"
            if val is not None:
                msg += f" | val {val(cnn)}"
            log(msg)
    return hist


def main():
    from .config import default_cfg
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--pool", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--lr-end", type=float, default=None)
    ap.add_argument("--scratch", action="store_true")
    ap.add_argument("--init", default=None)
    ap.add_argument("--out", required=True)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--crop-seeds", type=int, nargs="*", default=[10, 11, 12, 13])
    ap.add_argument("--crop-frac", type=float, default=0.5)
    args = ap.parse_args()
    cfg = default_cfg()
    ice = cfg.processes.iceflow
    nz = int(ice.numerics.Nz)
    dzeta = compute_zeta_dzeta(compute_levels(nz, float(ice.numerics.vert_spacing)))[1]
    dev = torch.device("cuda:0")
    weights = None
    if not args.scratch:
        init = args.init or os.path.join(EM.EMULATOR_DIR, "pinnbp_10_4_cnn_16_32_2_1_a.safetensors")
        weights, _ = EM.load_weights(init)
    cnn = EM.CNN(nz, int(ice.emulator.network.nb_layers), device=dev, weights=weights, seed=args.seed)
    # crops of 1000 x 1000 multi-dome fields drawn with other seeds than the
    # benchmark field (seed 2), which stays out of the training data
    crops = [syn.domes1000(seed=s) for s in args.crop_seeds]
    pool = make_pool(args.pool, seed=args.seed, dev=dev, crops_from=crops, crop_frac=args.crop_frac)
    train(cnn, pool, args.steps, dzeta, ice.physics, batch=args.batch, lr=args.lr, lr_end=args.lr_end,
          seed=args.seed)
    cnn.save(args.out, list(ice.emulator.fieldin), nz, float(ice.numerics.vert_spacing))
    print("saved", args.out)


if __name__ == "__main__":
    main()
