#!/usr/bin/env python3
"""Velocity fields/s at 1000 x 1000 x Nz=10: emulator vs energy solver at <= 1 % error.

One step = one velocity field of the emulated ice-flow update on the
synthetic 20-dome glacier (1000 x 1000, dx = 50 m, Nz = 10): CNN inference on
the MFMA kernels, then the shortest warm-started polish by the float32
energy solver (schedule searched here, before the timed region) that brings
the field to <= 1 % relative L2 error against the float64 minimiser, the
thk > 0 mask and derived fields, and the float64-reduced energy of the
fallback check.  Inputs are resident in HBM before the timed region.

Reported next to it:
* CNN inference alone (fields/s and its error);
* the float32 energy solver (the reference's plain method, legacy Adam)
  from a cold start, with the same schedule family and search, timed to
  <= 1 %, and the reference's own settings (lr = 1, no decay) as a curve;
* the float64 minimiser itself (float32 + float64 GPU solve, decaying step);
* live rooflines (HIP events on the launch stream) of the dominant kernel
  (the two-iteration energy/adjoint/Adam pass of the polish) and of the
  other hot kernels (one-iteration pass, energy + adjoint, 3x3 bf16 MFMA
  conv, transport);
* a CPU baseline: the oracle (torch-CPU float32 restatement) CNN inference
  on this host's cores.
"""
import argparse
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
# algorithmic HBM bytes per horizontal cell (Nz = 10, float32): U, V (80) and
# thk, usurf, arrhenius, slidingco (16) read; gU, gV (80) written -- and for
# the fused Adam iteration U, V, m, v of both read (240 + 16), written (240)
BYTES_ENERGY_ADJOINT = 176.0
BYTES_FUSED_ITERATION = 496.0
# transport: ubar, vbar, thk, smb, topg read (20), thk, usurf written (8)
BYTES_TRANSPORT = 28.0
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
# `bench.py --kernels-only`, scaled by a 1 GiB copy (the roofline's `traffic`)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_domes1000.json")
# schedule family searched for both paths: initial step, then total iterations
LR0S = (100.0, 300.0, 1000.0)
FLATS = (0.0, 0.1, 0.3)  # fraction of the iterations at the initial step
ENDS = (1.0, 10.0, 30.0, 60.0)  # step at the end of the geometric decay
# Adam (beta1, beta2): the Keras defaults the reference's solver uses, and two
# shorter-memory pairs (measured on this workload: from 410 to under 280
# iterations to 1 % from zero, and from 320 to about 230 from the emulator)
BETAS = ((0.9, 0.999), (0.95, 0.99), (0.95, 0.95))
CPU_SAMPLE_S = 15.0  # wall time of the CPU solver-iteration sample
REFINE_TOTALS = (100, 200, 300, 400, 500, 700, 1000, 1500, 2000, 3000, 4500, 7000, 10000, 14000)
SOLVER_TOTALS = (200, 300, 400, 500, 700, 1000, 1500, 2000, 3000, 4500, 7000, 10000, 14000, 20000)


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


def reference_solution(f32, f64, thk, nz, dzeta, n_flat=8000, n_decay=20000, n_refine=2000):
    """float64 minimiser of the workload (converges here: runs of 14k and 28k
    float32 iterations agree to 3e-3, and the float64 stage moves the field by
    ~1e-8)."""
    from igm_amd.energy import EnergyOperator
    from igm_amd.solver import converged_solution
    from oracle.energy import DEFAULT_PHYSICS
    op32 = EnergyOperator(dzeta, DEFAULT_PHYSICS, fields=f32)
    op64 = EnergyOperator(dzeta, DEFAULT_PHYSICS, fields=f64, dtype=torch.float64)
    z = torch.zeros((nz,) + tuple(thk.shape), dtype=torch.float32, device=thk.device)
    U, V, J, change = converged_solution(op32, op64, z, z.clone(), n_flat, n_decay, n_refine)
    return U, V, J, change, op64


def measured_traffic(kernel):
    """Counter-measured HBM bytes per launch of `kernel`, or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            return float(json.load(f)["kernels"][kernel]["traffic_bytes"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


def measure_kernels(cnn, op32, U, V, f32, vert_weight, nz, nlaunch=50):
    """Each hot kernel alone, back to back on the launch stream, timed with
    HIP events at the workload size: the 3x3 bf16 MFMA conv, the energy +
    adjoint pass, the fused Adam pass, the upwind transport, and a 1 GiB
    device copy for scale.  (U, V): a realistic velocity field."""
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
    # two solver iterations in one pass (moments read from one buffer, written to another)
    mom2 = torch.zeros_like(mom)
    part2 = torch.zeros_like(op32.partials)
    pair_us = timed(lambda: op32.run_pass2(U, V, Uo, Vo, mom, mom2, adam, adam, part2))
    # the upwind thickness update (superbee) on these velocities
    vo = velocity_outputs(U, V, thk, vert_weight, nz)
    smb = torch.zeros_like(thk)
    topg = (f32["usurf"] - thk).contiguous()
    thk_o, us_o = torch.empty_like(thk), torch.empty_like(thk)
    tr_us = timed(lambda: L.call("igm_transport_f32", Ny, Nx, L.ptr(vo["ubar"]), L.ptr(vo["vbar"]), L.ptr(thk),
                                 L.ptr(smb), L.ptr(topg), 50.0, 50.0, 0.01, 2, 0.91, 0.0, L.ptr(thk_o), None,
                                 L.ptr(us_o), None, st))
    # a plain device copy of 1 GiB, for scale (what HBM gives a streaming kernel)
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
         "pair_us": pair_us, "pair_gbs": BYTES_FUSED_ITERATION * cells / (pair_us * 1e-6) / 1e9,
         "transport_us": tr_us, "transport_gbs": BYTES_TRANSPORT * cells / (tr_us * 1e-6) / 1e9,
         "copy_gbs": copy_gbs}
    log(f"[bench] conv3x3 {conv_us:.1f} us/launch = {conv_tflops:.1f} TFLOP/s; energy+adjoint {en_us:.1f} us = "
        f"{r['energy_adjoint_gbs']:.0f} GB/s; fused Adam pass {fused_us:.1f} us = {r['fused_gbs']:.0f} GB/s; "
        f"two-iteration pass {pair_us:.1f} us ({pair_us / 2:.1f} us per iteration) = {r['pair_gbs']:.0f} GB/s; "
        f"transport {tr_us:.1f} us = {r['transport_gbs']:.0f} GB/s; device copy {copy_gbs:.0f} GB/s")
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ref-flat", type=int, default=8000)
    ap.add_argument("--ref-decay", type=int, default=20000)
    ap.add_argument("--solver-max", type=int, default=20000)
    ap.add_argument("--check-its", type=int, default=None, help="fallback check iterations (cfg default if unset)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--kernels-only", action="store_true",
                    help="time the hot kernels alone on the emulated field and exit (a short run for counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the ice-flow path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    from igm_amd import config, emulator as EM, iceflow, synthetic as syn
    from igm_amd import _lib as L
    from igm_amd.energy import EnergyOperator
    from igm_amd.fields import misfits, velocity_outputs
    from igm_amd.solver import LegacyAdamState, adam_schedule, rel_err, solve
    from oracle.energy import DEFAULT_PHYSICS

    dev = torch.device("cuda:0")
    L.load()
    nz = 10
    g = syn.domes1000()
    Ny, Nx = g["thk"].shape
    t0 = time.time()
    keys = ("thk", "usurf", "arrhenius", "slidingco", "dX")
    f32 = {k: torch.as_tensor(g[k], device=dev) for k in keys}
    f64 = {k: v.double() for k, v in f32.items()}
    thk = f32["thk"]
    mask = thk > 0
    cfg = config.default_cfg()
    ice = cfg.processes.iceflow
    ice.method = "emulated"
    ice.emulator.retrain_freq = 0
    if args.check_its is not None:
        ice.emulator.fallback.check_its = args.check_its

    # ------------------------------------------------------------ state --
    class State:
        pass

    state = State()
    for k in keys:
        setattr(state, k, f32[k])
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
    log(f"[bench] setup {time.time() - t0:.1f}s; weights {os.path.basename(wpath)}")

    if args.kernels_only:
        op32 = EnergyOperator(dzeta, DEFAULT_PHYSICS, fields=f32)
        EM.update_iceflow_emulated(cfg, state)
        kt = measure_kernels(state.iceflow_model, op32, state.U.contiguous(), state.V.contiguous(), f32,
                             state.vert_weight, nz, nlaunch=10)
        print(json.dumps({"kernels_us": {k: round(v, 2) for k, v in kt.items() if k.endswith("_us")}}), flush=True)
        return

    # ------------------------------------------------ float64 minimiser --
    t = time.time()
    Ur, Vr, J64, ref_change, op64 = reference_solution(f32, f64, thk, nz, dzeta, args.ref_flat, args.ref_decay)
    torch.cuda.synchronize()
    log(f"[bench] float64 reference: {args.ref_flat}+{args.ref_decay} its in {time.time() - t:.1f}s, "
        f"J={J64:.6f}, change under float64 refinement {ref_change:.2e}")

    op32 = EnergyOperator(dzeta, DEFAULT_PHYSICS, fields=f32)
    z = torch.zeros((nz, Ny, Nx), dtype=torch.float32, device=dev)

    def schedule_search(U0, V0, totals):
        """Shortest schedule of the family {Adam (beta1, beta2) in BETAS;
        lr0 in LR0S; a fraction in FLATS at lr0, then a geometric decay to a
        step in ENDS} from (U0, V0) that ends <= 1 %: the first total of
        ``totals`` that does, then bisected (to 2 iterations, one
        two-iteration pass) against the total before it.  Returns a dict of its, err, lr0, flat,
        lr_end, beta1, beta2 (the last total tried if none reaches 1 %)."""

        def trial(total):
            if total == 0:
                return {"its": 0, "err": rel_err(U0, V0, Ur, Vr, mask), "lr0": None, "flat": None,
                        "lr_end": None, "beta1": None, "beta2": None}
            cand = []
            for b1, b2 in BETAS:
                for lr0 in LR0S:
                    for ff in FLATS:
                        for lr_end in ENDS:
                            n_flat = int(ff * total)
                            r = solve(op32, U0, V0, total, LegacyAdamState(lr=lr0, beta1=b1, beta2=b2),
                                      early_stopping=False,
                                      lr_schedule=adam_schedule(lr0, n_flat, total - n_flat, lr_end))
                            cand.append({"its": total, "err": rel_err(r.U, r.V, Ur, Vr, mask), "lr0": lr0,
                                         "flat": ff, "lr_end": lr_end, "beta1": b1, "beta2": b2})
            best = min(cand, key=lambda c: c["err"])
            log(f"[bench]   {total} its: best rel err {best['err']:.4f} (lr0 {best['lr0']}, flat {best['flat']}, "
                f"end {best['lr_end']}, betas {best['beta1']}, {best['beta2']})")
            return best

        lo, best = None, None
        for total in totals:
            best = trial(total)
            if best["err"] <= 0.01:
                break
            lo = total
        if best["err"] > 0.01 or lo is None:
            return best
        hi = best["its"]
        for _ in range(8):
            mid = int(round((lo + hi) / 4.0)) * 2
            if mid <= lo or mid >= hi:
                break
            t = trial(mid)
            if t["err"] <= 0.01:
                hi, best = mid, t
            else:
                lo = mid
        return best

    # ------------------------------------------------- emulator path --
    # CNN inference alone, then the shortest warm-started polish that brings
    # the emulated field to <= 1 % (the timed emulated update runs both)
    EM.update_iceflow_emulated(cfg, state)
    U_emu, V_emu = state.U.clone(), state.V.clone()
    inf_err = rel_err(U_emu, V_emu, Ur, Vr, mask)
    vs = float(ice.numerics.vert_spacing)
    mis = {"emulator_only": misfits(thk, Ur, Vr, U_emu, V_emu, state.vert_weight, vs)}
    # the energy gap the fallback rule looks at, against the float64 minimiser
    j_emu = float(op32.energy(U_emu, V_emu)[0, 3])
    gap_emu = (j_emu - J64) / abs(J64)
    with Timer() as tm:
        for _ in range(args.steps):
            EM.update_iceflow_emulated(cfg, state)
    inf_ms = tm.ms / args.steps
    log(f"[bench] emulator inference {inf_ms:.3f} ms, rel err {inf_err:.4f}; searching the polish")
    pol = schedule_search(U_emu, V_emu, (0,) + REFINE_TOTALS)
    n_ref = pol["its"]
    ice.emulator.refine.its = int(n_ref)
    if n_ref > 0:
        ice.emulator.refine.lr = float(pol["lr0"])
        ice.emulator.refine.flat_frac = float(pol["flat"])
        ice.emulator.refine.lr_end = float(pol["lr_end"])
        ice.emulator.refine.beta1 = float(pol["beta1"])
        ice.emulator.refine.beta2 = float(pol["beta2"])

    def emulator_step():
        EM.update_iceflow_emulated(cfg, state)
        if ice.emulator.refine.its > 0:
            iceflow.refine_emulated(cfg, state)
        if ice.emulator.fallback.enabled:
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
    emu_err = rel_err(state.U, state.V, Ur, Vr, mask)
    mis["emulated"] = misfits(thk, Ur, Vr, state.U, state.V, state.vert_weight, vs)
    log(f"[bench] emulated update (inference + {n_ref} polish its) {emu_ms:.3f} ms ({emu_fps:.1f} fields/s), "
        f"rel err {emu_err:.4f}; fallbacks {state.fallback_count}/{args.steps}")

    # --------------------------------------------------- solver path --
    # (a) the reference's own settings: Adam lr = step_size = 1, no decay
    U, V = z, z.clone()
    opt = LegacyAdamState(lr=1.0)
    its, curve = 0, []
    for chunk in [1000] * 4 + [2000] * 3 + [4000] * 2:
        r = solve(op32, U, V, chunk, opt, early_stopping=False)
        U, V, its = r.U, r.V, its + chunk
        curve.append((its, round(rel_err(U, V, Ur, Vr, mask), 4)))
        if its >= args.solver_max:
            break
    ref_settings = {"lr": 1.0, "error_curve": curve}
    # (b) the same method and schedule family as the polish, from zero
    log("[bench] searching the solver schedule from zero")
    sol = schedule_search(z, z, SOLVER_TOTALS)
    n1, lr1, solver_err = sol["its"], sol["lr0"], sol["err"]
    solver_reached = solver_err <= 0.01
    n_flat1 = int(sol["flat"] * n1)
    sched = adam_schedule(lr1, n_flat1, n1 - n_flat1, sol["lr_end"])
    solve(op32, z, z.clone(), 20, LegacyAdamState(lr=1.0), early_stopping=False)
    with Timer() as tm:
        rs = solve(op32, z, z.clone(), n1, LegacyAdamState(lr=lr1, beta1=sol["beta1"], beta2=sol["beta2"]),
                   early_stopping=False, lr_schedule=sched)
    mis["solver"] = misfits(thk, Ur, Vr, rs.U, rs.V, state.vert_weight, vs)
    solver_ms = tm.ms
    solver_fps = 1000.0 / solver_ms
    it_ms = solver_ms / n1
    log(f"[bench] solver: {n1} its (lr0 {lr1}) to err {solver_err:.4f} (<= 1%: {solver_reached}) in "
        f"{solver_ms:.1f} ms ({it_ms * 1e3:.1f} us/it) -> {solver_fps:.3f} fields/s; reference settings {curve}")
    # reference early-stopping behaviour (nbitmax large)
    with Timer() as tm:
        res_es = solve(op32, z, z.clone(), 20000, LegacyAdamState(lr=1.0), early_stopping=True)
    es_ms = tm.ms
    es_err = rel_err(res_es.U, res_es.V, Ur, Vr, mask)
    log(f"[bench] reference settings with EarlyStopping: stops after {res_es.iterations} its "
        f"({es_ms:.1f} ms, {es_ms * 1e3 / max(1, res_es.iterations):.1f} us/it), rel err {es_err:.4f}")

    # ------------------------------------------------------- rooflines --
    kt = measure_kernels(state.iceflow_model, op32, Ur.float().contiguous(), Vr.float().contiguous(), f32,
                         state.vert_weight, nz)
    conv_us, conv_tflops = kt["conv_us"], kt["conv_tflops"]
    en_us, en_gbs = kt["energy_adjoint_us"], kt["energy_adjoint_gbs"]
    fused_us, fused_gbs = kt["fused_us"], kt["fused_gbs"]
    tr_us, tr_gbs, copy_gbs = kt["transport_us"], kt["transport_gbs"], kt["copy_gbs"]
    log(f"[bench] solver iteration (pass + finalize + gaps) {it_ms * 1e3:.1f} us; raw emulator energy gap "
        f"{gap_emu:.4%}")

    pair_us, pair_gbs = kt["pair_us"], kt["pair_gbs"]
    # the dominant kernel of the timed step: the two-iteration energy pass
    # when the field is polished (the solver runs two iterations per launch,
    # one leading single-iteration pass when the count is odd), the 3x3 conv
    # otherwise.  Its algorithmic bytes are one iteration's: the pass reads
    # and writes the state once for both.
    if n_ref > 0:
        traffic = measured_traffic("energy_pair_kernel<10> (two fused iterations)")
        roofline = {"bound": "hbm", "achieved": round(pair_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(pair_gbs / PEAK_HBM_GBS, 4),
                    "traffic": round(traffic) if traffic else None,
                    "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                    "kernel": "energy_pair_kernel<10> (two fused energy + adjoint + Adam iterations, 1000x1000)",
                    "bytes_per_cell": BYTES_FUSED_ITERATION, "us_per_launch": round(pair_us, 2),
                    "iterations_per_launch": 2, "launches_per_step": int(n_ref // 2)}
    else:
        traffic = measured_traffic("conv3x3_mfma_kernel<0> (32->32 bf16)")
        roofline = {"bound": "mfma", "achieved": round(conv_tflops, 2), "peak": PEAK_BF16_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(conv_tflops / PEAK_BF16_TFLOPS, 4),
                    "traffic": round(traffic) if traffic else None,
                    "kernel": "conv3x3_mfma_kernel<0> (32->32, 1000x1000)", "us_per_launch": round(conv_us, 2),
                    "launches_per_step": 15}

    # ----------------------------------------------------- CPU baseline --
    # the oracle (torch-CPU float32 restatement of the reference's TF path) on
    # this host: one emulator inference and a bounded number of solver
    # iterations (energy + autograd gradient + legacy Adam) on the same field,
    # scaled to the iteration counts the GPU paths need for <= 1 %
    cpu = None
    if not args.no_cpu:
        from oracle import cnn as OC
        from oracle import energy as OE
        from oracle.solver import LegacyAdam
        ncore = os.cpu_count() or 1
        torch.set_num_threads(min(16, ncore))
        X = torch.stack([torch.as_tensor(g[k]) for k in keys], -1)[None]
        weights = state.iceflow_model.get_weights()
        def median_time(fn, warm, n, budget_s):
            """Median wall time of >= n timed calls (fewer only past budget_s)
            after warm untimed ones."""
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
        fin = OE.fieldin_tensors(*(g[k] for k in keys))
        Uc = torch.zeros((nz, Ny, Nx))
        Vc = torch.zeros_like(Uc)
        mvs = [torch.zeros_like(Uc) for _ in range(4)]
        opt = LegacyAdam(lr=float(lr1))

        def cpu_iteration():
            _, _, gU_, gV_ = OE.energy_and_grad(Uc, Vc, fin, dzeta, DEFAULT_PHYSICS)
            opt.apply([Uc, Vc], [gU_, gV_], [mvs[0], mvs[2]], [mvs[1], mvs[3]])

        cpu_it_s, k = median_time(cpu_iteration, 2, 5, CPU_SAMPLE_S)
        cpu_emu_s = cpu_inf_s + n_ref * cpu_it_s
        cpu = {"value": round(1.0 / cpu_emu_s, 8), "unit": "velocity fields/s", "cores": torch.get_num_threads(),
               "kind": "port",
               "sample": f"oracle emulator inference of the 1000x1000 field (median of {n_inf}, {cpu_inf_s:.2f} s) "
                         f"+ oracle solver iterations (median of {k} after 2 warm-ups, {cpu_it_s:.2f} s each), "
                         f"scaled to the {n_ref} polish iterations of the GPU emulated path; TensorFlow is not "
                         "installed",
               "solver_fields_per_s": round(1.0 / (n1 * cpu_it_s), 8),
               "emulator_inference_fields_per_s": round(1.0 / cpu_inf_s, 4),
               "iteration_s": round(cpu_it_s, 4)}
        log(f"[bench] CPU oracle: CNN inference {cpu_inf_s:.2f} s ({n_inf} timed), solver iteration "
            f"{cpu_it_s:.2f} s ({k} timed) on {torch.get_num_threads()} threads")

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
        "config": {"workload": "domes1000: 1000x1000 grid, dx=50 m, Nz=10, 20 seeded domes on a random-field bed; "
                               "emulated iceflow update (bf16 MFMA CNN 16x3x3x32 + 1x1, fp32 first layer) "
                               "polished by the shortest warm-started fp32 energy-solver schedule that reaches "
                               "<= 1 % rel err vs the float64 minimiser, plus the float64 energy-gap check",
                   "weights": os.path.basename(wpath),
                   "polish_its": int(n_ref), "polish_lr0": pol["lr0"], "polish_flat_frac": pol["flat"],
                   "polish_lr_end": pol["lr_end"], "polish_adam_betas": [pol["beta1"], pol["beta2"]],
                   "fallback_check_its": int(ice.emulator.fallback.check_its)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "emulated_rel_err_vs_fp64": round(emu_err, 5),
        # the reference's own accuracy measures against the float64 minimiser:
        # pretraining.py:502-535 (rel L1/L2, thickness-weighted layer
        # midpoints) and diagnostic.py:74-84 (velbar L1 / RMS, m/yr)
        "reference_misfits": {k: {n: round(x, 6) for n, x in v.items()} for k, v in mis.items()},
        "emulator_inference": {"ms": round(inf_ms, 4), "fields_per_s": round(1000.0 / inf_ms, 1),
                               "rel_err_vs_fp64": round(inf_err, 5), "energy_gap_vs_fp64": round(gap_emu, 6)},
        "fallbacks": int(state.fallback_count),
        "solver_fields_per_s": round(solver_fps, 4),
        "solver_schedule": {"lr0": lr1, "flat": int(n_flat1), "decay": int(n1 - n_flat1), "lr_end": sol["lr_end"],
                            "adam_betas": [sol["beta1"], sol["beta2"]],
                            "rel_err": round(solver_err, 5)},
        "solver_iters_to_1pct": int(n1),
        "solver_reached_1pct": bool(solver_reached),
        "solver_reference_settings": ref_settings,
        "solver_us_per_iter": round(it_ms * 1e3, 2),
        "emulator_vs_solver_speedup": round(emu_fps / solver_fps, 1),
        "solver_early_stopping": {"iterations": int(res_es.iterations), "rel_err": round(es_err, 5),
                                  "ms": round(es_ms, 2)},
        "roofline_fused_single": {"bound": "hbm", "achieved": round(fused_gbs, 1), "peak": PEAK_HBM_GBS,
                                  "unit": "GB/s", "frac": round(fused_gbs / PEAK_HBM_GBS, 4),
                                  "kernel": "energy_pass_kernel<float, 10, 2, 0> (one fused iteration)",
                                  "traffic": measured_traffic(
                                      "energy_pass_kernel<float,10,2,0> (fused energy + adjoint + Adam)"),
                                  "bytes_per_cell": BYTES_FUSED_ITERATION, "us_per_launch": round(fused_us, 2)},
        "roofline_energy_adjoint": {"bound": "hbm", "achieved": round(en_gbs, 1), "peak": PEAK_HBM_GBS,
                                    "unit": "GB/s", "frac": round(en_gbs / PEAK_HBM_GBS, 4),
                                    "bytes_per_cell": BYTES_ENERGY_ADJOINT, "us_per_launch": round(en_us, 2)},
        "roofline_conv3x3": {"bound": "mfma", "achieved": round(conv_tflops, 2), "peak": PEAK_BF16_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(conv_tflops / PEAK_BF16_TFLOPS, 4),
                             "us_per_launch": round(conv_us, 2),
                             "hbm_gbs": round(2 * 64.0 * Ny * Nx / (conv_us * 1e-6) / 1e9, 1)},
        "roofline_transport": {"bound": "hbm", "achieved": round(tr_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": round(tr_gbs / PEAK_HBM_GBS, 4), "bytes_per_cell": BYTES_TRANSPORT,
                               "us_per_launch": round(tr_us, 2)},
        "solver_iteration_us": round(it_ms * 1e3, 2),
        "device_copy_gbs": round(copy_gbs, 1),
        "fp64_reference": {"J": J64, "its": args.ref_flat + args.ref_decay, "change_under_refinement": ref_change},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
