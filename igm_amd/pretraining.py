"""Physics-informed emulator training on synthetic glaciers.

The analogue of igm/processes/pretraining/pretraining.py:198-493
(train_iceflow_emulator) with generated data instead of netCDF datasets:
each step draws a batch of glacier patches of uneven sizes, packs them into
one zero-padded NHWC batch with per-entry extents (the kernels mask
everything outside an entry's extent), and takes one plain legacy-Adam step
(no clipnorm, pretraining.py:222-228) on the mean over the batch of the
Blatter-Pattyn energy of the predicted field (pretraining.py:318-338: loss =
energy, no labels; the reference's batches hold equal-size samples, where
its reduce_mean over all cells is this mean of per-entry means).
Augmentation: random flips and transposition of each patch
(pretraining.py:550-559).

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
    """Batches of ``batch`` pool patches (random picks, random flips and
    transposition).  ``buckets`` > 1 sorts a batch by patch area and splits
    it into that many sub-batches, each zero-padded only to its own largest
    patch: the gradients of the sub-batches add up to the one of the whole
    batch (padding is masked), with less padded work."""

    def __init__(self, pool, batch, seed, dev, buckets=1):
        self.pool, self.batch, self.dev = pool, batch, dev
        self.buckets = max(1, min(int(buckets), batch))
        self.rng = np.random.default_rng(seed)
        self.pixels = self.padded = 0  # valid and padded pixels served so far
        # the 8 flip / transpose variants of every patch, made once and kept
        # contiguous (8x the pool's memory; a step then only copies)
        self.variants = [[augment(p, c).contiguous() for c in range(8)] for p in pool]

    def _pack(self, parts):
        H = max(p.shape[0] for p in parts)
        W = max(p.shape[1] for p in parts)
        self.pixels += sum(p.shape[0] * p.shape[1] for p in parts)
        self.padded += len(parts) * H * W
        X = torch.zeros((len(parts), H, W, 5), dtype=torch.float32, device=self.dev)
        ext = torch.empty((len(parts), 2), dtype=torch.int32, pin_memory=True)
        for b, p in enumerate(parts):
            X[b, :p.shape[0], :p.shape[1]] = p
            ext[b, 0], ext[b, 1] = p.shape[0], p.shape[1]
        # pinned host buffer + non-blocking copy: no host/device synchronisation per step
        return X, ext.to(self.dev, non_blocking=True)

    def next(self):
        """List of (X (b, H, W, 5), extents (b, 2) int32) sub-batches."""
        idx = self.rng.integers(0, len(self.pool), self.batch)
        codes = self.rng.integers(0, 8, self.batch)
        parts = [self.variants[i][int(c)] for i, c in zip(idx, codes)]
        if self.buckets == 1:
            return [self._pack(parts)]
        parts.sort(key=lambda p: p.shape[0] * p.shape[1])
        cuts = np.linspace(0, len(parts), self.buckets + 1).round().astype(int)
        return [self._pack(parts[c0:c1]) for c0, c1 in zip(cuts[:-1], cuts[1:]) if c1 > c0]


class PatchBatchBatcher:
    """Fresh batches: step s serves ``synthetic.patch_batch(s)`` -- 32 patches
    of sides 128..512 seeded by 3 + s (numpy on the host) -- as one
    zero-padded batch with per-entry extents (``buckets`` = 1: one launch
    per layer over the whole batch, the padding masked) or split by area
    into ``buckets`` padded sub-batches."""

    def __init__(self, dev, batch=32, start=0, buckets=1):
        self.dev, self.batch, self.step, self.buckets = dev, batch, start, max(1, int(buckets))
        self.pixels = self.padded = 0

    def next(self):
        parts = [patch_tensor(p, self.dev) for p in syn.patch_batch(self.step, n=self.batch)]
        self.step += 1
        return self.pack(parts)

    def pack(self, parts):
        """(X, extents) sub-batches of a list of (h, w, 5) patches."""
        packer = Batcher.__new__(Batcher)
        packer.dev, packer.pixels, packer.padded = self.dev, 0, 0
        if self.buckets == 1:
            out = [packer._pack(parts)]
        else:
            parts = sorted(parts, key=lambda p: p.shape[0] * p.shape[1])
            cuts = np.linspace(0, len(parts), self.buckets + 1).round().astype(int)
            out = [packer._pack(parts[c0:c1]) for c0, c1 in zip(cuts[:-1], cuts[1:]) if c1 > c0]
        self.pixels += packer.pixels
        self.padded += packer.padded
        return out


class CropBatcher:
    """Batches of ``batch`` square crops of side ``crop`` from whole fields
    (random field, offset, flips and transposition): equal-size entries, so
    no padding and no extents."""

    def __init__(self, fields, batch, crop, seed, dev):
        self.fields, self.batch, self.crop, self.dev = fields, batch, int(crop), dev
        for f in fields:
            assert f.shape[0] >= self.crop and f.shape[1] >= self.crop, "crop larger than a field"
        self.rng = np.random.default_rng(seed)
        self.pixels = self.padded = 0

    def next(self):
        c = self.crop
        self.pixels += self.batch * c * c
        self.padded += self.batch * c * c
        X = torch.empty((self.batch, c, c, 5), dtype=torch.float32, device=self.dev)
        for b in range(self.batch):
            f = self.fields[int(self.rng.integers(len(self.fields)))]
            j0 = int(self.rng.integers(0, f.shape[0] - c + 1))
            i0 = int(self.rng.integers(0, f.shape[1] - c + 1))
            X[b] = augment(f[j0:j0 + c, i0:i0 + c], int(self.rng.integers(0, 8)))
        return [(X, None)]


class Validator:
    """Relative velocity error of the emulator on one whole field against
    the field's converged float64 energy minimiser (solver.converged_solution)."""

    def __init__(self, field, dzeta, phys, dev, n_flat=8000, n_decay=20000):
        from .solver import converged_solution
        f32 = {k: torch.as_tensor(field[k], device=dev) for k in KEYS}
        f64 = {k: v.double() for k, v in f32.items()}
        op32 = EnergyOperator(dzeta, phys, fields=f32)
        op64 = EnergyOperator(dzeta, phys, fields=f64, dtype=torch.float64)
        self.nz = len(dzeta) + 1
        z = torch.zeros((self.nz,) + tuple(f32["thk"].shape), dtype=torch.float32, device=dev)
        self.Ur, self.Vr, self.J, self.change = converged_solution(op32, op64, z, z.clone(), n_flat, n_decay)
        self.mask = f32["thk"] > 0
        self._keep = [f32[k] for k in KEYS]
        self.fields = EM.planar_fields(self._keep)

    def __call__(self, cnn):
        from .solver import rel_err
        Y = cnn.forward(self.fields)
        U = Y[0, ..., :self.nz].permute(2, 0, 1)
        V = Y[0, ..., self.nz:].permute(2, 0, 1)
        return rel_err(U, V, self.Ur, self.Vr, self.mask)


class WeightAverage:
    """Exponential moving average of the flat fp32 weights (Polyak
    averaging): ``swap`` exchanges it with the live weights (and repacks the
    bf16 operands), so validation and saving can use the average."""

    def __init__(self, cnn, decay):
        self.cnn, self.decay = cnn, float(decay)
        self.w = cnn.w.clone()

    def update(self):
        self.w.mul_(self.decay).add_(self.cnn.w, alpha=1.0 - self.decay)

    def swap(self):
        live = self.cnn.w.clone()
        self.cnn.w.copy_(self.w)
        self.w.copy_(live)
        self.cnn.pack()


def train(cnn, batcher, dzeta, phys, steps=None, seconds=None, lr=1e-4, lr_end=None, clipnorm=None,
          epsilon=1e-7, log_every=100, val=None, on_best=None, log=print, ema=None):
    """Energy-as-loss training for ``steps`` steps or ``seconds`` of wall
    time (whichever is given); returns the list of (step, mean COST per
    entry[, validation error]).  The loss of a step is the mean over the
    batch's entries of their COST; ``clipnorm`` (per-tensor clip_by_norm,
    as the emulator's online retraining uses) is off unless given.

    lr decays geometrically from ``lr`` to ``lr_end`` over the run (constant
    if None).  ``val(cnn)`` (optional) is evaluated every log_every steps;
    ``on_best(cnn, v)`` is called when it improves.  With ``ema`` (a decay,
    e.g. 0.999) the validation and ``on_best`` see the weight average."""
    assert (steps is None) != (seconds is None), "give steps or seconds"
    hist = []
    best = float("inf")
    avg = WeightAverage(cnn, ema) if ema else None
    t0 = time.time()
    s = 0
    while True:
        frac = (s / max(1, steps - 1)) if steps is not None else (time.time() - t0) / seconds
        if (steps is not None and s >= steps) or (seconds is not None and frac >= 1.0):
            break
        costs = []
        parts = batcher.next()
        n_entries = sum(int(X.shape[0]) for X, _ in parts)
        for k, (X, ext) in enumerate(parts):
            op = EnergyOperator(dzeta, phys, X=X, extents=ext)
            fields = EM.nhwc_fields(X)
            Y, acts = cnn.forward(fields, extents=ext, keep=True)
            e, gY = op.energy_grad(Y)
            gY.mul_(1.0 / n_entries)  # d(mean over the batch of COST_b)/dY
            cnn.backward(gY, acts, fields, extents=ext, accumulate=k > 0, deferred=True)
            costs.append(e[:, 3])
        cnn.reduce_deferred()
        lr_s = lr if lr_end is None else lr * (lr_end / lr) ** min(1.0, frac)
        cnn.adam_step(lr_s, clipnorm=0.0 if clipnorm is None else clipnorm, epsilon=epsilon)
        if avg is not None:
            avg.update()
        s += 1
        if s % log_every == 0 or s == 1:
            c = float(torch.cat(costs).mean())
            rec = [s, c]
            msg = f"step {s}: mean COST {c:.5f} lr {lr_s:.2e} ({time.time() - t0:.1f}s)"
            if val is not None:
                if avg is not None:
                    avg.swap()
                v = float(val(cnn))
                rec.append(v)
                msg += f" | val {v:.4f}"
                if v < best:
                    best = v
                    if on_best is not None:
                        on_best(cnn, v)
                if avg is not None:
                    avg.swap()
            hist.append(tuple(rec))
            log(msg)
    torch.cuda.synchronize()
    dt = time.time() - t0
    px, pad = getattr(batcher, "pixels", 0), getattr(batcher, "padded", 0)
    log(f"trained {s} steps in {dt:.1f}s: {dt / max(1, s) * 1e3:.2f} ms/step, {px / dt:.4g} patch pixels/s "
        f"({pad / dt:.4g} with the padding, {pad / max(1, px):.2f}x)")
    return hist


def main():
    """``python -m igm_amd.pretraining --out <file> [...]``.

    --data patches: glacier_patch() draws of 128..512 cells plus crops of
    multi-dome fields (the uneven-size batches of pretraining.py);
    --data domes: square crops of whole 1000 x 1000 multi-dome fields drawn
    with --field-seeds.  --val-seed S validates on the domes1000 field of
    seed S against its converged float64 solution and keeps the best
    weights; otherwise the last weights are saved."""
    from .config import default_cfg
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--seconds", type=float, default=None)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--data", choices=["patches", "domes"], default="patches")
    ap.add_argument("--pool", type=int, default=256)
    ap.add_argument("--buckets", type=int, default=1)
    ap.add_argument("--crop", type=int, default=512)
    ap.add_argument("--field-seeds", type=int, nargs="*", default=list(range(100, 132)))
    ap.add_argument("--lr", type=float, default=1e-4)
    ap.add_argument("--lr-end", type=float, default=None)
    ap.add_argument("--scratch", action="store_true")
    ap.add_argument("--init", default=None)
    ap.add_argument("--out", required=True)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--crop-seeds", type=int, nargs="*", default=[10, 11, 12, 13])
    ap.add_argument("--crop-frac", type=float, default=0.5)
    ap.add_argument("--val-seed", type=int, default=None)
    ap.add_argument("--log-every", type=int, default=100)
    ap.add_argument("--ema", type=float, default=None, help="validate/save an EMA of the weights (decay)")
    args = ap.parse_args()
    if args.steps is None and args.seconds is None:
        args.steps = 2000
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
    t0 = time.time()
    if args.data == "domes":
        fields = [patch_tensor(syn.domes1000(seed=s), dev) for s in args.field_seeds]
        batcher = CropBatcher(fields, args.batch, args.crop, args.seed, dev)
    else:
        # crops of 1000 x 1000 multi-dome fields drawn with seeds of their own
        crops = [syn.domes1000(seed=s) for s in args.crop_seeds]
        pool = make_pool(args.pool, seed=args.seed, dev=dev, crops_from=crops, crop_frac=args.crop_frac)
        batcher = Batcher(pool, args.batch, args.seed, dev, buckets=args.buckets)
    print(f"data ready in {time.time() - t0:.1f}s", flush=True)
    save = lambda m: m.save(args.out, list(ice.emulator.fieldin), nz, float(ice.numerics.vert_spacing))
    val = on_best = None
    if args.val_seed is not None:
        t0 = time.time()
        val = Validator(syn.domes1000(seed=args.val_seed), dzeta, ice.physics, dev)
        print(f"validation field ready in {time.time() - t0:.1f}s (float64 stage change {val.change:.2e}); "
              f"initial error {val(cnn):.4f}", flush=True)
        on_best = lambda m, v: save(m)
    hist = train(cnn, batcher, dzeta, ice.physics, steps=args.steps, seconds=args.seconds, lr=args.lr,
                 lr_end=args.lr_end, log_every=args.log_every, val=val, on_best=on_best,
                 log=lambda m: print(m, flush=True), ema=args.ema)
    if val is None:
        save(cnn)
    else:
        print(f"best validation error {min(h[2] for h in hist):.4f}", flush=True)
    print("saved", args.out, flush=True)


if __name__ == "__main__":
    main()
