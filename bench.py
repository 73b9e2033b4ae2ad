#!/usr/bin/env python3
"""Velocity fields/s at 1000 x 1000 x Nz=10: emulator vs energy solver at <= 1 % error.

One step = one velocity field of the emulated ice-flow update on the
synthetic 20-dome glacier (1000 x 1000, dx = 50 m, Nz = 10, seed 2): CNN
inference on the MFMA kernels, a warm-started polish by the float32 energy
solver, the thk > 0 mask and derived fields, and the energy-gap fallback
check (solver iterations from the polished field; a gap above 1 % would
fall back to a full solve).  Inputs are resident in HBM before the timed
region.

The polish schedule, and the cold-start solver's, are chosen before the
timed region on held-out fields (domes1000 seeds 21 and 22, never used for
training): the shortest schedule of one family that brings every held-out
field to <= 0.9 % relative error against its float64 minimiser -- and, for
the polish, whose polished held-out fields also pass the timed step's
energy-gap check with a gap <= 0.8 %.  The benchmark field's own error at
that fixed length is what is reported.

Reported next to it:
* CNN inference alone (fields/s and its error);
* the float32 energy solver (the reference's plain method, legacy Adam)
  from a cold start, schedule chosen the same way, timed on the benchmark
  field, and the reference's own settings (lr = 1, no decay) as a curve;
* the float64 minimiser and its certificate (a run of twice the length,
  the float64 refinement, the gradient norm);
* live rooflines (HIP events on the launch stream) of the dominant kernel
  (the two-iteration energy/adjoint/Adam pass of the polish) and of the
  other hot kernels (one-iteration pass, energy + adjoint, 3x3 bf16 MFMA
  conv, transport);
* emulator training on fresh batches of 32 patches of sides 128..512;
* the 1500 x 1500 transient run (emulated vs solved);
* a CPU baseline: the oracle (torch-CPU float32 restatement) on this host.
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "velocity fields/s at 1000²×Nz10: emulator vs energy solver at ≤1% rel err"
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0
BENCH_WEIGHTS = ("igm_amd_cnn_16_32_nz10.safetensors", "pinnbp_10_4_cnn_16_32_2_1_a.safetensors")
KEYS = ("thk", "usurf", "arrhenius", "slidingco", "dX")
# algorithmic HBM bytes per horizontal cell (Nz = 10, float32): U, V (80) and
# thk, usurf, arrhenius, slidingco (16) read; gU, gV (80) written -- and for
# one fused Adam iteration U, V, m, v of both read (240 + 16), written (240)
BYTES_ENERGY_ADJOINT = 176.0
BYTES_FUSED_ITERATION = 496.0
# transport: ubar, vbar, thk, smb, topg read (20), thk, usurf written (8)
BYTES_TRANSPORT = 28.0
# emulator FLOPs per pixel: 2 (9*5*32 + 15*9*32*32 + 32*20)
FLOP_PER_PIXEL = 280640.0
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
# `bench.py --kernels-only`, scaled by a 1 GiB copy (the roofline's `traffic`)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_domes1000.json")
# schedule family searched for both paths: Adam (beta1, beta2), initial step,
# fraction of the iterations at it, step at the end of a geometric decay, and
# a geometric warm-up to the initial step over a fraction of the iterations
# (none, or 10 % from lr0 / 6).  The polish from the emulated field wants small
# steps (lr0 ~30), the cold start large ones (~150): the family spans both.
BETAS = ((0.93, 0.95), (0.95, 0.95), (0.95, 0.98))
LR0S = (20.0, 30.0, 45.0, 70.0, 100.0, 150.0)
FLATS = (0.1, 0.2, 0.3, 0.5)
ENDS = (2.0, 5.0, 10.0, 20.0, 30.0)
GRID_RAMPS = ((0.0, 1.0), (0.1, 1.0 / 6.0))
# Adam's epsilon: below the per-node gradients of a 1000 x 1000 field (~1e-8 ..
# 1e-6: the energy is a mean over 10^6 cells), where the Keras default 1e-7
# turns Adam into plain gradient descent on small ones (1e-7 won no length of
# either leg once both were searched)
EPSS = (1e-9,)
# warm-up variants tried around the two best grid schedules of each length: a
# geometric ramp over a fraction of the iterations, from lr0 * r up to lr0 or
# to WARM_PEAK * lr0 (the warm-up lets a larger step follow)
RAMPS = (0.05, 0.1, 0.15)
RAMP_FROM = (1.0 / 6.0, 0.25, 0.4)
WARM_PEAK = 4.0 / 3.0
LOCAL_ROUNDS = 3  # rounds of coordinate moves around the best schedule of a length
TUNE_SEEDS = (21, 22)     # held-out domes1000 fields the schedules are chosen on
TUNE_TARGET = 0.009       # every held-out field must reach this error
# a polished held-out field must also pass the energy-gap check with this
# margin under its threshold (1 %): J_emu - J_check <= GAP_TARGET |J_check|
GAP_TARGET = 0.008
REFINE_TOTALS = (80, 100, 120, 150, 200, 250, 300, 400, 500, 700, 1000, 1500, 2000, 3000)
SOLVER_TOTALS = (150, 200, 250, 300, 400, 500, 700, 1000, 1500, 2000, 3000, 4500)
CPU_SAMPLE_S = 15.0  # wall time of the CPU solver-iteration sample


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    """HIP events on torch's current stream (the stream every kernel here is launched on)."""

    def __init__(self):
        self.a = torch.cuda.Event(enable_timing=True)
        self.b = torch.cuda.Event(enable_timing=True)

    def __enter__(self):
        self.a.record()
        return self

    def __exit__(self, *exc):
        self.b.record()
        self.b.synchronize()
        self.ms = self.a.elapsed_time(self.b)


def measured_traffic(kernel):
    """Counter-measured HBM bytes per launch of `kernel`, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            return float(json.load(f)["kernels"][kernel]["traffic_bytes"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class Field:
    """One domes1000 field on the device: its float32/float64 energy
    operators, two float64 minimisers from independent schedules and the
    emulator's field.  The minimiser's velocities are weakly determined in
    thin ice (runs of different schedules end ~1e-3 apart at energies equal
    to ~1e-7), so an error is measured against both and the larger counts;
    their distance is the reference's uncertainty."""

    def __init__(self, seed, phys, dzeta, cnn, nz, dev, n_flat=8000, n_decay=20000):
        from igm_amd import synthetic as syn
        from igm_amd.energy import EnergyOperator
        from igm_amd.solver import converged_solution
        self.seed = seed
        g = syn.domes1000(seed=seed)
        self.g = g
        self.f32 = {k: torch.as_tensor(g[k], device=dev) for k in KEYS}
        f64 = {k: v.double() for k, v in self.f32.items()}
        self.op32 = EnergyOperator(dzeta, phys, fields=self.f32)
        self.op64 = EnergyOperator(dzeta, phys, fields=f64, dtype=torch.float64)
        self.mask = self.f32["thk"] > 0
        z = torch.zeros((nz,) + tuple(g["thk"].shape), dtype=torch.float32, device=dev)
        self.z = z
        # A: 28,000 float32 + 2,000 float64 iterations, Keras epsilon;
        # B: 56,000 + 2,000, epsilon 1e-14, decay to 1e-4
        self.Ur, self.Vr, self.J, self.change = converged_solution(self.op32, self.op64, z, z.clone(), n_flat,
                                                                   n_decay)
        self.Ur_b, self.Vr_b, self.J_b, _ = converged_solution(self.op32, self.op64, z, z.clone(), 2 * n_flat,
                                                               2 * n_decay, lr_end=1e-4, epsilon=1e-14)
        from igm_amd.solver import rel_err
        self.uncertainty = rel_err(self.Ur, self.Vr, self.Ur_b, self.Vr_b, self.mask)
        from igm_amd.emulator import planar_fields
        Y = cnn.forward(planar_fields([self.f32[k] for k in KEYS]))
        self.Ue = torch.where(self.mask, Y[0, ..., :nz].permute(2, 0, 1), 0.0).contiguous()
        self.Ve = torch.where(self.mask, Y[0, ..., nz:].permute(2, 0, 1), 0.0).contiguous()

    def err(self, U, V):
        """Relative L2 error over the ice against the two minimisers (the larger)."""
        from igm_amd.solver import rel_err
        return max(rel_err(U, V, self.Ur, self.Vr, self.mask), rel_err(U, V, self.Ur_b, self.Vr_b, self.mask))


def run_schedule(op, U0, V0, sch):
    from igm_amd.solver import LegacyAdamState, adam_schedule, solve
    n = sch["its"]
    n_warm = int(sch.get("ramp", 0.0) * n)
    n_flat = int(sch["flat"] * n)
    return solve(op, U0, V0, n, LegacyAdamState(lr=sch["lr0"], beta1=sch["beta1"], beta2=sch["beta2"],
                                                epsilon=sch["eps"]),
                 early_stopping=False,
                 lr_schedule=adam_schedule(sch["lr0"], n_flat, n - n_warm - n_flat, sch["lr_end"], n_warmup=n_warm,
                                           lr_start=sch.get("lr_start", sch["lr0"])))


def warmup_variants(sch):
    """The warm-up neighbourhood of a grid schedule: a geometric ramp over a
    fraction in RAMPS of the iterations from a step lr0 * r, r in RAMP_FROM,
    up to lr0 or WARM_PEAK * lr0, then its flat part and its decay."""
    out = []
    for peak in (1.0, WARM_PEAK):
        for ramp in RAMPS:
            for r in RAMP_FROM:
                lr0 = sch["lr0"] * peak
                out.append(dict(sch, lr0=lr0, ramp=ramp, lr_start=lr0 * r))
    return out


def check_gap(fd, U, V, check_its, lr, counters):
    """The energy-gap check of the fallback on a polished field (U, V) of
    field ``fd``, as the emulated step runs it: nodes with no ice within one
    node zeroed, ``check_its`` solver iterations with the solver's step
    ``lr`` and a legacy-Adam step counter at each of ``counters`` (the check
    keeps one optimizer, whose counter grows by check_its per call); returns
    the largest (J_emu - J_check) / |J_check|."""
    from igm_amd.solver import LegacyAdamState, solve
    near = getattr(fd, "near_ice", None)
    if near is None:
        ice = (fd.f32["thk"] > 0).float()[None, None]
        near = fd.near_ice = torch.nn.functional.max_pool2d(ice, 3, stride=1, padding=1)[0, 0] > 0
    U0 = torch.where(near, U, 0.0).contiguous()
    V0 = torch.where(near, V, 0.0).contiguous()
    gaps = []
    for t in counters:
        opt = LegacyAdamState(lr=lr)
        opt.iterations = int(t)
        c = solve(fd.op32, U0, V0, check_its, opt, early_stopping=False).costs
        j_emu = float(c[0])
        j_chk = min(float(np.min(c)), j_emu)
        gaps.append((j_emu - j_chk) / abs(j_chk))
    return max(gaps)


def local_variants(sch):
    """Coordinate moves around a schedule: lr0 (with its warm-up start)
    scaled by 0.8 .. 1.25, the flat fraction +-0.1, the end step scaled by
    0.5 .. 2 (not above lr0), beta1 and beta2 +-0.02."""
    out = [dict(sch, lr0=sch["lr0"] * f, lr_start=sch["lr_start"] * f) for f in (0.8, 0.9, 1.1, 1.25)]
    out += [dict(sch, flat=round(ff, 3)) for ff in (sch["flat"] - 0.1, sch["flat"] + 0.1) if 0.0 <= ff <= 0.8]
    out += [dict(sch, lr_end=sch["lr_end"] * f) for f in (0.5, 0.7, 1.4, 2.0) if sch["lr_end"] * f <= sch["lr0"]]
    out += [dict(sch, beta1=round(b, 4)) for b in (sch["beta1"] - 0.02, sch["beta1"] + 0.02) if 0.0 < b < 1.0]
    out += [dict(sch, beta2=round(b, 4)) for b in (sch["beta2"] - 0.02, sch["beta2"] + 0.02) if 0.0 < b < 1.0]
    return out


def schedule_search(fields, start, totals, target, gap=None, gap_target=None):
    """Shortest schedule that brings EVERY field of ``fields`` (from
    ``start(field)``) to an error <= target.  At each trial length: the grid
    {Adam (beta1, beta2) in BETAS and epsilon in EPSS; a warm-up in
    GRID_RAMPS; lr0 in LR0S; a fraction in FLATS at lr0, then a geometric
    decay to a step in ENDS (not above lr0)}, then
    the warm-up variants (``warmup_variants``) of its two best schedules,
    then rounds of coordinate moves (``local_variants``) around the best
    while it improves; the best is the lowest worst-case error.  With ``gap`` (a function of the
    field and the schedule's result), a schedule counts only if it is <=
    ``gap_target`` on every field.  The length is the first of ``totals``
    whose best schedule meets the target, then bisected to 2 iterations
    against the total before it.  Returns the schedule dict with the
    worst-case error ``err``, the per-field errors ``errs`` and, with
    ``gap``, the per-field ``gaps``."""
    none = {"err": float("inf")}

    def evaluate(sch, bound):
        errs, outs = [], []
        for fd in fields:
            U0, V0 = start(fd)
            r = run_schedule(fd.op32, U0, V0, sch)
            errs.append(fd.err(r.U, r.V))
            outs.append((fd, r.U, r.V))
            if max(errs) >= bound:  # cannot enter the two best
                return None
        gaps = [gap(*o) for o in outs] if gap is not None else []
        if gaps and max(gaps) > gap_target:
            return None
        return dict(sch, err=max(errs), errs=errs, gaps=gaps)

    def trial(total):
        top = []  # the two best, lowest worst-case error first

        def consider(sch):
            bound = top[-1]["err"] if len(top) == 2 else float("inf")
            r = evaluate(sch, bound)
            if r is not None:
                top.append(r)
                top.sort(key=lambda x: x["err"])
                del top[2:]

        for eps, (b1, b2), lr0, ff, lr_end, (ramp, r0) in itertools.product(EPSS, BETAS, LR0S, FLATS, ENDS,
                                                                            GRID_RAMPS):
            if lr_end <= lr0:
                consider({"its": total, "lr0": lr0, "flat": ff, "lr_end": lr_end, "beta1": b1, "beta2": b2,
                          "eps": eps, "ramp": ramp, "lr_start": lr0 * r0})
        for base in list(top):
            for sch in warmup_variants(base):
                consider(sch)
        # then up to LOCAL_ROUNDS rounds of coordinate moves around the best
        for _ in range(LOCAL_ROUNDS if top else 0):
            base = top[0]
            for sch in local_variants(base):
                consider(sch)
            if top[0] is base:
                break
        if not top:
            log(f"[bench]   {total} its: no schedule passes the energy-gap check on every held-out field")
            return dict(none, its=total)
        best = top[0]
        gaps = f", worst gap {max(best['gaps']):.2e}" if best["gaps"] else ""
        log(f"[bench]   {total} its: worst held-out err {best['err']:.4f}{gaps} (lr0 {best['lr0']:g}, warm-up "
            f"{best['ramp']:g} from {best['lr_start']:g}, flat {best['flat']}, end {best['lr_end']}, betas "
            f"{best['beta1']}, {best['beta2']}, eps {best['eps']:g})")
        return best

    lo, best = None, None
    for total in totals:
        best = trial(total)
        if best["err"] <= target:
            break
        lo = total
    if "lr0" not in best:
        raise RuntimeError(f"no schedule of up to {totals[-1]} iterations passes the energy-gap check")
    if best["err"] > target or lo is None:
        return best
    hi = best["its"]
    for _ in range(8):
        mid = int(round((lo + hi) / 4.0)) * 2
        if mid <= lo or mid >= hi:
            break
        t = trial(mid)
        if t["err"] <= target:
            hi, best = mid, t
        else:
            lo = mid
    return best


def measure_kernels(cnn, op32, U, V, f32, vert_weight, nz, nlaunch=50):
    """Each hot kernel alone, back to back on the launch stream, timed with
    HIP events at the workload size: the 3x3 bf16 MFMA conv, the energy +
    adjoint pass, the fused Adam pass, the two-iteration pass, the upwind
    transport, and a 1 GiB device copy for scale.  (U, V): a realistic
    velocity field."""
    from igm_amd import _lib as L
    from igm_amd.fields import velocity_outputs
    thk = f32["thk"]
    Ny, Nx = thk.shape
    dev = thk.device
    st = L.stream_ptr()
    a_in = torch.zeros((1, Ny, Nx, 32), dtype=torch.bfloat16, device=dev)
    a_in.copy_(torch.randn((1, Ny, Nx, 32), device=dev).to(torch.bfloat16))
    a_out = torch.empty_like(a_in)

    def timed(fn):
        for _ in range(5):
            fn()
        with Timer() as tm:
            for _ in range(nlaunch):
                fn()
        return tm.ms * 1e3 / nlaunch

    conv_us = timed(lambda: L.call("igm_conv3x3_fwd_bf16", L.ptr(a_in), L.ptr(cnn.fwd[0]), L.ptr(cnn.bias(1)),
                                   None, 1, Ny, Nx, L.ptr(a_out), st))
    conv_tflops = 2.0 * 9 * 32 * 32 * Ny * Nx / (conv_us * 1e-6) / 1e12
    gU, gV = torch.empty_like(U), torch.empty_like(V)
    en_us = timed(lambda: op32.run_pass(1, U, V, gU=gU, gV=gV))
    mom = torch.zeros((nz, Ny, Nx, 4), dtype=torch.float32, device=dev)
    Uo, Vo = torch.empty_like(U), torch.empty_like(V)
    adam = L.IgmAdam(alpha=1e-3, beta1=0.9, beta2=0.999, epsilon=1e-7)
    fused_us = timed(lambda: op32.run_pass(2, U, V, Uo=Uo, Vo=Vo, moments=mom, adam=adam))
    mom2 = torch.zeros_like(mom)
    part2 = torch.zeros_like(op32.partials)
    # the two-iteration pass as a fixed-count solve runs it between its first
    # and last launch: interleaved (U, V) in and out
    UVi = torch.stack([U, V], dim=-1).contiguous()
    UVo = torch.empty_like(UVi)
    pair_us = timed(lambda: op32.run_pass2(UVi, None, UVo, None, mom, mom2, adam, adam, part2))
    vo = velocity_outputs(U, V, thk, vert_weight, nz)
    smb = torch.zeros_like(thk)
    topg = (f32["usurf"] - thk).contiguous()
    thk_o, us_o = torch.empty_like(thk), torch.empty_like(thk)
    tr_us = timed(lambda: L.call("igm_transport_f32", Ny, Nx, L.ptr(vo["ubar"]), L.ptr(vo["vbar"]), L.ptr(thk),
                                 L.ptr(smb), L.ptr(topg), 50.0, 50.0, 0.01, 2, 0.91, 0.0, L.ptr(thk_o), None,
                                 L.ptr(us_o), None, st))
    ca = torch.empty(2 ** 28, dtype=torch.float32, device=dev).fill_(1.0)
    cb = torch.empty_like(ca)
    for _ in range(3):
        cb.copy_(ca)
    with Timer() as tm:
        for _ in range(10):
            cb.copy_(ca)
    copy_gbs = 2 * ca.numel() * 4 * 10 / (tm.ms * 1e-3) / 1e9
    del ca, cb
    cells = Ny * Nx
    r = {"conv_us": conv_us, "conv_tflops": conv_tflops,
         "energy_adjoint_us": en_us, "energy_adjoint_gbs": BYTES_ENERGY_ADJOINT * cells / (en_us * 1e-6) / 1e9,
         "fused_us": fused_us, "fused_gbs": BYTES_FUSED_ITERATION * cells / (fused_us * 1e-6) / 1e9,
         "pair_us": pair_us, "pair_gbs": 2 * BYTES_FUSED_ITERATION * cells / (pair_us * 1e-6) / 1e9,
         "transport_us": tr_us, "transport_gbs": BYTES_TRANSPORT * cells / (tr_us * 1e-6) / 1e9,
         "copy_gbs": copy_gbs}
    log(f"[bench] conv3x3 {conv_us:.1f} us/launch = {conv_tflops:.1f} TFLOP/s; energy+adjoint {en_us:.1f} us = "
        f"{r['energy_adjoint_gbs']:.0f} GB/s; fused Adam pass {fused_us:.1f} us = {r['fused_gbs']:.0f} GB/s; "
        f"two-iteration pass {pair_us:.1f} us ({pair_us / 2:.1f} us per iteration) = {r['pair_gbs']:.0f} GB/s "
        f"algorithmic; transport {tr_us:.1f} us = {r['transport_gbs']:.0f} GB/s; device copy {copy_gbs:.0f} GB/s")
    return r


def bench_training(cnn_weights, nz, dzeta, phys, dev, parts_per_step, buckets):
    """Emulator training on fresh batches: step s trains on patch_batch(s)
    (32 glacier patches of sides 128..512, seeded by 3 + s), generated and
    moved to the GPU before the timed region (``parts_per_step``: the
    patches of each step; the first two steps are warm-up).  Returns ms per
    step, patch pixels/s and the algorithmic convolution TFLOP/s (forward +
    both backward products: 3 x 280,640 FLOP per valid pixel)."""
    from igm_amd import emulator as EM
    from igm_amd import pretraining as PT
    src = PT.PatchBatchBatcher(dev, start=0, buckets=buckets)
    batches = [src.pack(parts) for parts in parts_per_step]
    steps = len(batches) - 2
    pixels = src.pixels / len(batches)

    class Replay:
        def __init__(self, bs):
            self.bs, self.i = bs, 0

        def next(self):
            b = self.bs[self.i % len(self.bs)]
            self.i += 1
            return b

    cnn = EM.CNN(nz, 16, device=dev, weights=cnn_weights)
    PT.train(cnn, Replay(batches[:2]), dzeta, phys, steps=2, lr=1e-4, log_every=10 ** 9, log=lambda m: None)
    torch.cuda.synchronize()
    rep = Replay(batches[2:])
    with Timer() as tm:
        PT.train(cnn, rep, dzeta, phys, steps=steps, lr=1e-4, log_every=10 ** 9, log=lambda m: None)
    ms = tm.ms / steps
    tflops = 3 * FLOP_PER_PIXEL * pixels / (ms * 1e-3) / 1e12
    return {"ms_per_step": round(ms, 3), "patch_pixels_per_s": round(pixels / (ms * 1e-3), 0),
            "sub_batches": buckets, "conv_tflops_algorithmic": round(tflops, 1),
            "conv_frac_of_bf16_peak": round(tflops / PEAK_BF16_TFLOPS, 4),
            "padded_over_valid": round(src.padded / max(1, src.pixels), 3)}


def bench_transient(weights_path, years):
    """The 1500 x 1500 transient run (smb_simple -> iceflow -> time -> thk,
    CFL 0.3, step_max 1 yr, SMB = min(0.005 (usurf - 2800), 2) m/yr),
    emulated (online retraining every 10 steps, the energy-gap check every
    10th update) and solved (the reference's plain method)."""
    from igm_amd import synthetic as syn
    from igm_amd import transient as TR
    g = syn.domes1500()
    out = {"workload": "domes1500: 1500x1500, dx=50 m, 40 domes, seed 4", "years": years}
    for method in ("emulated", "solved"):
        cfg = TR.transient_cfg(method=method, years=years)
        cfg.processes.iceflow.emulator.name = weights_path
        st = TR.make_state(g)
        t0 = time.perf_counter()
        rec = TR.run(cfg, st)
        wall = time.perf_counter() - t0
        v0 = float(rec["volume"][0])
        keep = [t for t in TR.REPORT_YEARS if t <= years and np.interp(t, rec["t"], rec["volume"]) >= 0.1 * v0]
        out[method] = {"steps": int(len(rec["t"])), "iceflow_fields_per_s": round(1000.0 / float(np.mean(rec["iceflow_ms"])), 2),
                       "iceflow_ms_median": round(float(np.median(rec["iceflow_ms"])), 3),
                       "fallbacks": int(rec["fallbacks"][-1]), "wall_s": round(wall, 1),
                       "initial_volume_km3": round(v0, 3),
                       "volume_km3_at": {f"{t:g}": round(float(np.interp(t, rec["t"], rec["volume"])), 4) for t in keep}}
    ve, vs = out["emulated"]["volume_km3_at"], out["solved"]["volume_km3_at"]
    out["volume_rel_diff_at"] = {t: round(abs(ve[t] - vs[t]) / vs[t], 6) for t in vs if t in ve and vs[t] > 0}
    out["compared_while"] = "volume >= 10 % of the initial volume in both runs"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--solver-max", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the training and transient workloads")
    ap.add_argument("--train-steps", type=int, default=20)
    ap.add_argument("--transient-years", type=float, default=200.0)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--kernels-only", action="store_true",
                    help="time the hot kernels alone on the emulated field and exit (a short run for counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the ice-flow path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    from igm_amd import config, emulator as EM, iceflow
    from igm_amd import _lib as L
    from igm_amd.fields import misfits
    from igm_amd.solver import LegacyAdamState, adam_schedule, converged_solution, rel_err, solve

    dev = torch.device("cuda:0")
    L.load()
    nz = 10
    t0 = time.time()
    cfg = config.default_cfg()
    ice = cfg.processes.iceflow
    ice.method = "emulated"
    ice.emulator.retrain_freq = 0
    phys = ice.physics

    # ------------------------------------------------------------ state --
    from igm_amd import synthetic as syn
    g = syn.domes1000(seed=2)

    class State:
        pass

    state = State()
    for k in KEYS:
        setattr(state, k, torch.as_tensor(g[k], device=dev))
    state.it = -1
    wpath = args.weights
    if wpath is None:
        for w in BENCH_WEIGHTS:
            p = os.path.join(EM.EMULATOR_DIR, w)
            if os.path.exists(p):
                wpath = p
                break
    ice.emulator.pretrained = True
    ice.emulator.name = wpath
    ice.emulator.nbit_init = 0  # no retraining inside the bench
    iceflow.initialize(cfg, state)
    dzeta = np.asarray(state.dzeta)
    cnn = state.iceflow_model
    log(f"[bench] setup {time.time() - t0:.1f}s; weights {os.path.basename(wpath)}")

    if args.kernels_only:
        from igm_amd.energy import EnergyOperator
        op32 = EnergyOperator(dzeta, phys, fields={k: getattr(state, k) for k in KEYS})
        EM.update_iceflow_emulated(cfg, state)
        kt = measure_kernels(cnn, op32, state.U.contiguous(), state.V.contiguous(),
                             {k: getattr(state, k) for k in KEYS}, state.vert_weight, nz, nlaunch=10)
        print(json.dumps({"kernels_us": {k: round(v, 2) for k, v in kt.items() if k.endswith("_us")}}), flush=True)
        return

    # ----------------------------------- float64 minimisers + certificate --
    t = time.time()
    bench_f = Field(2, phys, dzeta, cnn, nz, dev)
    tune_f = [Field(s, phys, dzeta, cnn, nz, dev) for s in TUNE_SEEDS]
    # certificate of the benchmark field's minimisers: their distance, and
    # the float64 gradient norm at A against the one at zero
    cert_change = bench_f.uncertainty
    _, gU, gV = bench_f.op64.energy_grad(bench_f.Ur.contiguous(), bench_f.Vr.contiguous())
    z64 = bench_f.z.double()
    _, g0U, g0V = bench_f.op64.energy_grad(z64, z64.clone())
    grad_ratio = float(torch.linalg.vector_norm(torch.stack([gU, gV])) /
                       torch.linalg.vector_norm(torch.stack([g0U, g0V])))
    torch.cuda.synchronize()
    log(f"[bench] float64 minimisers (two schedules each) of seeds 2, {TUNE_SEEDS} in {time.time() - t:.1f}s; "
        f"seed 2: J_A={bench_f.J:.9f}, J_B={bench_f.J_b:.9f}, distance A-B {cert_change:.2e} (held-out: "
        f"{', '.join(f'{f.uncertainty:.2e}' for f in tune_f)}), |grad|/|grad(0)| at A {grad_ratio:.2e}")
    Ur, Vr, mask = bench_f.Ur, bench_f.Vr, bench_f.mask
    thk = state.thk
    op32 = bench_f.op32

    # ------------------------------------------------- emulator path --
    EM.update_iceflow_emulated(cfg, state)
    U_emu, V_emu = state.U.clone(), state.V.clone()
    inf_err = bench_f.err(U_emu, V_emu)
    vs = float(ice.numerics.vert_spacing)
    mis = {"emulator_only": misfits(thk, Ur, Vr, U_emu, V_emu, state.vert_weight, vs)}
    j_emu = float(op32.energy(U_emu, V_emu)[0, 3])
    gap_emu = (j_emu - bench_f.J) / abs(bench_f.J)
    with Timer() as tm:
        for _ in range(args.steps):
            EM.update_iceflow_emulated(cfg, state)
    inf_ms = tm.ms / args.steps
    log(f"[bench] emulator inference {inf_ms:.3f} ms, rel err {inf_err:.4f} (held-out seeds: "
        f"{', '.join(f'{f.err(f.Ue, f.Ve):.4f}' for f in tune_f)}); choosing the polish on seeds {TUNE_SEEDS}")
    fb = ice.emulator.fallback
    # the check's optimizer counter over the timed steps (it starts at zero
    # and grows by check_its per check): its first and its last value
    n_chk = int(fb.check_its)
    counters = (n_chk * args.warmup, n_chk * (args.warmup + args.steps - 1))
    pol = schedule_search(tune_f, lambda f: (f.Ue, f.Ve), REFINE_TOTALS, TUNE_TARGET,
                          gap=lambda f, U, V: check_gap(f, U, V, n_chk, float(ice.solver.step_size), counters),
                          gap_target=GAP_TARGET)
    rf = ice.emulator.refine
    rf.its, rf.lr, rf.flat_frac, rf.lr_end, rf.beta1, rf.beta2, rf.epsilon = (
        int(pol["its"]), float(pol["lr0"]), float(pol["flat"]), float(pol["lr_end"]), float(pol["beta1"]),
        float(pol["beta2"]), float(pol["eps"]))
    rf.warmup_frac, rf.warmup_lr = float(pol["ramp"]), float(pol["lr_start"])

    def emulator_step():
        EM.update_iceflow_emulated(cfg, state)
        iceflow.refine_emulated(cfg, state)
        iceflow.fallback_check(cfg, state)

    for _ in range(args.warmup):
        emulator_step()
    torch.cuda.synchronize()
    state.fallback_count = 0
    state.fallback_calls = 0
    with Timer() as tm:
        for _ in range(args.steps):
            emulator_step()
    emu_ms = tm.ms / args.steps
    emu_fps = 1000.0 / emu_ms
    emu_err = bench_f.err(state.U, state.V)
    gap_check = (float(state.COST_EMULATED) - state.COST_CHECK) / abs(state.COST_CHECK)
    mis["emulated"] = misfits(thk, Ur, Vr, state.U, state.V, state.vert_weight, vs)
    log(f"[bench] emulated update (inference + {rf.its} polish its + {fb.check_its}-iteration check) {emu_ms:.3f} ms "
        f"({emu_fps:.1f} fields/s), rel err {emu_err:.4f}; fallbacks {state.fallback_count}/{args.steps}, "
        f"energy gap at the check {gap_check:.2e}")

    # --------------------------------------------------- solver path --
    z = bench_f.z
    U, V = z, z.clone()
    opt = LegacyAdamState(lr=1.0)
    its, curve = 0, []
    for chunk in [1000] * 4 + [2000] * 3 + [4000] * 2:
        r = solve(op32, U, V, chunk, opt, early_stopping=False)
        U, V, its = r.U, r.V, its + chunk
        curve.append((its, round(bench_f.err(U, V), 4)))
        if its >= args.solver_max:
            break
    log(f"[bench] choosing the cold-start solver schedule on seeds {TUNE_SEEDS}")
    sol = schedule_search(tune_f, lambda f: (f.z, f.z), SOLVER_TOTALS, TUNE_TARGET)
    run_schedule(op32, z, z.clone(), dict(sol, its=20))
    with Timer() as tm:
        rs = run_schedule(op32, z, z.clone(), sol)
    solver_ms = tm.ms
    solver_err = bench_f.err(rs.U, rs.V)
    mis["solver"] = misfits(thk, Ur, Vr, rs.U, rs.V, state.vert_weight, vs)
    solver_fps = 1000.0 / solver_ms
    it_ms = solver_ms / sol["its"]
    log(f"[bench] solver: {sol['its']} its to err {solver_err:.4f} in {solver_ms:.1f} ms ({it_ms * 1e3:.1f} us/it) "
        f"-> {solver_fps:.3f} fields/s; reference settings {curve}")
    with Timer() as tm:
        res_es = solve(op32, z, z.clone(), 20000, LegacyAdamState(lr=1.0), early_stopping=True)
    es_ms = tm.ms
    es_err = bench_f.err(res_es.U, res_es.V)
    log(f"[bench] reference settings with EarlyStopping: stops after {res_es.iterations} its ({es_ms:.1f} ms), "
        f"rel err {es_err:.4f}")

    # ------------------------------------------------------- rooflines --
    kt = measure_kernels(cnn, op32, Ur.float().contiguous(), Vr.float().contiguous(), bench_f.f32,
                         state.vert_weight, nz)
    cells = float(thk.numel())
    traffic = measured_traffic("energy_pair_kernel<10, 3> (two fused iterations, interleaved U, V)")
    roofline = {"bound": "hbm", "achieved": round(kt["pair_gbs"], 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(kt["pair_gbs"] / PEAK_HBM_GBS, 4),
                "traffic": round(traffic) if traffic else None,
                "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                "traffic_gbs": round(traffic / (kt["pair_us"] * 1e-6) / 1e9, 1) if traffic else None,
                "traffic_frac": round(traffic / (kt["pair_us"] * 1e-6) / 1e9 / PEAK_HBM_GBS, 4) if traffic else None,
                "kernel": "energy_pair_kernel<10, 3> (two fused energy + adjoint + Adam iterations, interleaved "
                          "(U, V) in and out, 1000x1000)",
                "algorithmic_bytes_per_launch": 2 * BYTES_FUSED_ITERATION * cells,
                "bytes_per_cell_per_iteration": BYTES_FUSED_ITERATION, "iterations_per_launch": 2,
                "us_per_launch": round(kt["pair_us"], 2), "launches_per_step": int((rf.its + 1) // 2)}

    extras = {}
    if not args.no_extras:
        t = time.time()
        w, _ = EM.load_weights(wpath)
        from igm_amd.pretraining import patch_tensor
        parts = [[patch_tensor(p, dev) for p in syn.patch_batch(s)] for s in range(args.train_steps + 2)]
        extras["training_patch_batch"] = {
            "workload": "patch_batch(s): 32 glacier patches of sides 128..512 seeded by 3 + s, a fresh batch per "
                        "step, energy-as-loss Adam step (lr 1e-4)", "steps": args.train_steps,
            "one_launch_per_layer": bench_training(w, nz, dzeta, phys, dev, parts, 1),
            "four_sub_batches": bench_training(w, nz, dzeta, phys, dev, parts, 4)}
        del parts
        log(f"[bench] training on fresh batches ({time.time() - t:.1f}s): {extras['training_patch_batch']}")
        t = time.time()
        extras["transient"] = bench_transient(wpath, args.transient_years)
        log(f"[bench] transient 1500x1500 ({time.time() - t:.1f}s): {extras['transient']}")

    # ----------------------------------------------------- CPU baseline --
    cpu = None
    if not args.no_cpu:
        from oracle import cnn as OC
        from oracle import energy as OE
        from oracle.solver import LegacyAdam
        ncore = os.cpu_count() or 1
        torch.set_num_threads(min(16, ncore))
        X = torch.stack([torch.as_tensor(g[k]) for k in KEYS], -1)[None]
        weights = cnn.get_weights()
        pdict = {k: getattr(phys, k) for k in ("exp_glen", "exp_weertman", "regu_glen", "regu_weertman",
                                               "thr_ice_thk", "min_sr", "max_sr", "ice_density", "gravity_cst",
                                               "force_negative_gravitational_energy")}

        def median_time(fn, warm, n, budget_s):
            for _ in range(warm):
                fn()
            ts, t_all = [], time.time()
            while len(ts) < n and (len(ts) < 2 or time.time() - t_all < budget_s):
                t = time.time()
                fn()
                ts.append(time.time() - t)
            return float(np.median(ts)), len(ts)

        with torch.no_grad():
            cpu_inf_s, n_inf = median_time(lambda: OC.forward(X, weights), 2, 5, CPU_SAMPLE_S / 3)
        fin = OE.fieldin_tensors(*(g[k] for k in KEYS))
        Uc = torch.zeros((nz,) + tuple(g["thk"].shape))
        Vc = torch.zeros_like(Uc)
        mvs = [torch.zeros_like(Uc) for _ in range(4)]
        copt = LegacyAdam(lr=float(sol["lr0"]))

        def cpu_iteration():
            _, _, gU_, gV_ = OE.energy_and_grad(Uc, Vc, fin, dzeta, pdict)
            copt.apply([Uc, Vc], [gU_, gV_], [mvs[0], mvs[2]], [mvs[1], mvs[3]])

        cpu_it_s, k = median_time(cpu_iteration, 2, 5, CPU_SAMPLE_S)
        n_emu_its = rf.its + int(fb.check_its)
        cpu_emu_s = cpu_inf_s + n_emu_its * cpu_it_s
        cpu = {"value": round(1.0 / cpu_emu_s, 8), "unit": "velocity fields/s", "cores": torch.get_num_threads(),
               "cpu_model": cpu_model(), "kind": "port",
               "sample": f"oracle emulator inference of the 1000x1000 field (median of {n_inf}, {cpu_inf_s:.2f} s) "
                         f"+ oracle solver iterations (median of {k} after 2 warm-ups, {cpu_it_s:.2f} s each), "
                         f"scaled to the {n_emu_its} solver iterations of the GPU emulated step (polish + check); "
                         "TensorFlow is not installed",
               "solver_fields_per_s": round(1.0 / (sol["its"] * cpu_it_s), 8),
               "emulator_inference_fields_per_s": round(1.0 / cpu_inf_s, 4),
               "iteration_s": round(cpu_it_s, 4)}
        log(f"[bench] CPU oracle ({cpu['cpu_model']}): CNN inference {cpu_inf_s:.2f} s ({n_inf} timed), solver "
            f"iteration {cpu_it_s:.2f} s ({k} timed) on {torch.get_num_threads()} threads")

    def sched_out(s):
        return {"its": int(s["its"]), "lr0": s["lr0"], "warmup_frac": s["ramp"], "warmup_from": s["lr_start"],
                "flat_frac": s["flat"], "lr_end": s["lr_end"],
                "adam_betas": [s["beta1"], s["beta2"]], "adam_epsilon": s["eps"],
                "held_out_errs": [round(e, 5) for e in s["errs"]],
                **({"held_out_energy_gaps": [round(g, 5) for g in s["gaps"]]} if s.get("gaps") else {})}

    out = {
        "metric": METRIC,
        "value": round(emu_fps, 3),
        "unit": "velocity fields/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(emu_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16+fp32",  # bf16 MFMA convolutions, fp32 energy/adjoint/Adam polish
        "data": "synthetic",
        "config": {"workload": "domes1000: 1000x1000 grid, dx=50 m, Nz=10, 20 seeded domes on a random-field bed "
                               "(seed 2); emulated iceflow update (bf16 MFMA CNN 16x3x3x32 + 1x1, fp32 first "
                               "layer) polished by the fp32 energy solver with a schedule chosen on held-out "
                               f"fields (seeds {list(TUNE_SEEDS)}, <= {TUNE_TARGET:.1%} on each, energy gap at "
                               f"the check <= {GAP_TARGET:.1%}), then the energy-gap fallback check",
                   "weights": os.path.basename(wpath),
                   "polish": sched_out(pol),
                   "fallback_check_its": int(fb.check_its), "fallback_rel_energy_gap": float(fb.rel_energy_gap)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "emulated_rel_err_vs_fp64": round(emu_err, 5),
        "emulated_reached_1pct": bool(emu_err <= 0.01),
        "reference_misfits": {k: {n: round(x, 6) for n, x in v.items()} for k, v in mis.items()},
        "emulator_inference": {"ms": round(inf_ms, 4), "fields_per_s": round(1000.0 / inf_ms, 1),
                               "rel_err_vs_fp64": round(inf_err, 5), "energy_gap_vs_fp64": round(gap_emu, 6),
                               "held_out_rel_errs": [round(f.err(f.Ue, f.Ve), 5) for f in tune_f]},
        "fallbacks": int(state.fallback_count),
        "fallback_energy_gap": gap_check,
        "solver_fields_per_s": round(solver_fps, 4),
        "solver_schedule": sched_out(sol),
        "solver_rel_err_vs_fp64": round(solver_err, 5),
        "solver_reached_1pct": bool(solver_err <= 0.01),
        "solver_reference_settings": {"lr": 1.0, "error_curve": curve},
        "solver_us_per_iter": round(it_ms * 1e3, 2),
        "emulator_vs_solver_speedup": round(emu_fps / solver_fps, 3),
        "solver_early_stopping": {"iterations": int(res_es.iterations), "rel_err": round(es_err, 5),
                                  "ms": round(es_ms, 2)},
        "roofline_fused_single": {"bound": "hbm", "achieved": round(kt["fused_gbs"], 1), "peak": PEAK_HBM_GBS,
                                  "unit": "GB/s", "frac": round(kt["fused_gbs"] / PEAK_HBM_GBS, 4),
                                  "kernel": "energy_pass_kernel<float, 10, 2, 0> (one fused iteration)",
                                  "traffic": measured_traffic(
                                      "energy_pass_kernel<float,10,2,0> (fused energy + adjoint + Adam)"),
                                  "bytes_per_cell": BYTES_FUSED_ITERATION, "us_per_launch": round(kt["fused_us"], 2)},
        "roofline_energy_adjoint": {"bound": "hbm", "achieved": round(kt["energy_adjoint_gbs"], 1),
                                    "peak": PEAK_HBM_GBS, "unit": "GB/s",
                                    "frac": round(kt["energy_adjoint_gbs"] / PEAK_HBM_GBS, 4),
                                    "bytes_per_cell": BYTES_ENERGY_ADJOINT,
                                    "us_per_launch": round(kt["energy_adjoint_us"], 2)},
        # a 32->32 3x3 layer does 18,432 FLOP per 128 B moved (bf16 in and out):
        # 144 FLOP/B, below the bf16 ridge (2.5 PFLOP/s / 8 TB/s = 312), so the
        # roofline bounds it by HBM; the MFMA fraction is reported beside it
        "roofline_conv3x3": {"bound": "hbm", "achieved": round(2 * 64.0 * cells / (kt["conv_us"] * 1e-6) / 1e9, 1),
                             "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": round(2 * 64.0 * cells / (kt["conv_us"] * 1e-6) / 1e9 / PEAK_HBM_GBS, 4),
                             "bytes_per_pixel": 128.0, "flop_per_byte": 144.0,
                             "ridge_flop_per_byte": round(PEAK_BF16_TFLOPS * 1e3 / PEAK_HBM_GBS, 1),
                             "mfma_tflops": round(kt["conv_tflops"], 2),
                             "mfma_frac": round(kt["conv_tflops"] / PEAK_BF16_TFLOPS, 4),
                             "us_per_launch": round(kt["conv_us"], 2)},
        "roofline_transport": {"bound": "hbm", "achieved": round(kt["transport_gbs"], 1), "peak": PEAK_HBM_GBS,
                               "unit": "GB/s", "frac": round(kt["transport_gbs"] / PEAK_HBM_GBS, 4),
                               "bytes_per_cell": BYTES_TRANSPORT, "us_per_launch": round(kt["transport_us"], 2)},
        "device_copy_gbs": round(kt["copy_gbs"], 1),
        "fp64_reference": {"minimisers": "A: 28,000 float32 + 2,000 float64 Adam iterations (epsilon 1e-7); "
                                         "B: 56,000 + 2,000 (epsilon 1e-14); errors are the larger of the two",
                           "J_A": bench_f.J, "J_B": bench_f.J_b, "distance_A_B": cert_change,
                           "float64_refinement_change_A": bench_f.change, "grad_norm_over_grad_at_zero_A": grad_ratio,
                           "held_out": {str(f.seed): {"J_A": f.J, "J_B": f.J_b, "distance_A_B": f.uncertainty}
                                        for f in tune_f}},
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
