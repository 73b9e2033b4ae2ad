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
  (the 3x3 bf16 MFMA conv) and of the energy/adjoint kernel;
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
# schedule family searched for both paths: initial step, then total iterations
LR0S = (3.0, 10.0, 30.0)
REFINE_TOTALS = (100, 200, 300, 500, 700, 1000, 1500, 2000, 3000, 4500, 7000, 10000, 14000)
SOLVER_TOTALS = (1000, 1500, 2000, 3000, 4500, 7000, 10000, 14000, 20000)


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
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the ice-flow path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    from igm_amd import config, emulator as EM, iceflow, synthetic as syn
    from igm_amd import _lib as L
    from igm_amd.energy import EnergyOperator
    from igm_amd.fields import velocity_outputs
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

    # ------------------------------------------------ float64 minimiser --
    t = time.time()
    Ur, Vr, J64, ref_change, op64 = reference_solution(f32, f64, thk, nz, dzeta, args.ref_flat, args.ref_decay)
    torch.cuda.synchronize()
    log(f"[bench] float64 reference: {args.ref_flat}+{args.ref_decay} its in {time.time() - t:.1f}s, "
        f"J={J64:.6f}, change under float64 refinement {ref_change:.2e}")

    op32 = EnergyOperator(dzeta, DEFAULT_PHYSICS, fields=f32)
    z = torch.zeros((nz, Ny, Nx), dtype=torch.float32, device=dev)

    def schedule_search(U0, V0, totals):
        """Shortest schedule of the family {lr0 in LR0S; 30 % flat, then a
        geometric decay to 1e-3} from (U0, V0) that ends <= 1 %: returns
        (its, lr0, err) of the first total that does (or of the last tried)."""
        best = None
        for total in totals:
            if total == 0:
                cand = [(0, None, rel_err(U0, V0, Ur, Vr, mask))]
            else:
                cand = []
                for lr0 in LR0S:
                    n_flat = int(0.3 * total)
                    r = solve(op32, U0, V0, total, LegacyAdamState(lr=lr0), early_stopping=False,
                              lr_schedule=adam_schedule(lr0, n_flat, total - n_flat))
                    cand.append((total, lr0, rel_err(r.U, r.V, Ur, Vr, mask)))
            best = min(cand, key=lambda c: c[2])
            log(f"[bench]   {total} its: best rel err {best[2]:.4f} (lr0 {best[1]})")
            if best[2] <= 0.01:
                break
        return best

    # ------------------------------------------------- emulator path --
    # CNN inference alone, then the shortest warm-started polish that brings
    # the emulated field to <= 1 % (the timed emulated update runs both)
    EM.update_iceflow_emulated(cfg, state)
    U_emu, V_emu = state.U.clone(), state.V.clone()
    inf_err = rel_err(U_emu, V_emu, Ur, Vr, mask)
    with Timer() as tm:
        for _ in range(args.steps):
            EM.update_iceflow_emulated(cfg, state)
    inf_ms = tm.ms / args.steps
    log(f"[bench] emulator inference {inf_ms:.3f} ms, rel err {inf_err:.4f}; searching the polish")
    n_ref, lr_ref, _ = schedule_search(U_emu, V_emu, (0,) + REFINE_TOTALS)
    ice.emulator.refine.its = int(n_ref)
    if lr_ref is not None:
        ice.emulator.refine.lr = float(lr_ref)

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
    n1, lr1, solver_err = schedule_search(z, z, SOLVER_TOTALS)
    solver_reached = solver_err <= 0.01
    n_flat1 = int(0.3 * n1)
    sched = adam_schedule(lr1, n_flat1, n1 - n_flat1)
    solve(op32, z, z.clone(), 20, LegacyAdamState(lr=1.0), early_stopping=False)
    with Timer() as tm:
        rs = solve(op32, z, z.clone(), n1, LegacyAdamState(lr=lr1), early_stopping=False, lr_schedule=sched)
    solver_ms = tm.ms
    solver_fps = 1000.0 / solver_ms
    it_ms = solver_ms / n1
    log(f"[bench] solver: {n1} its (lr0 {lr1}) to err {solver_err:.4f} (<= 1%: {solver_reached}) in "
        f"{solver_ms:.1f} ms ({it_ms * 1e3:.1f} us/it) -> {solver_fps:.3f} fields/s; reference settings {curve}")
    # reference early-stopping behaviour (nbitmax large)
    res_es = solve(op32, z, z.clone(), 20000, LegacyAdamState(lr=1.0), early_stopping=True)
    es_err = rel_err(res_es.U, res_es.V, Ur, Vr, mask)

    # ------------------------------------------------------- rooflines --
    cnn = state.iceflow_model
    fields = EM.planar_fields([state.thk, state.usurf, state.arrhenius, state.slidingco, state.dX])
    a_in = torch.zeros((1, Ny, Nx, 32), dtype=torch.bfloat16, device=dev)
    a_in.copy_(torch.randn((1, Ny, Nx, 32), device=dev).to(torch.bfloat16))
    a_out = torch.empty_like(a_in)
    nlaunch = 50
    st = L.stream_ptr()
    for _ in range(5):
        L.call("igm_conv3x3_fwd_bf16", L.ptr(a_in), L.ptr(cnn.fwd[0]), L.ptr(cnn.bias(1)), None, 1, Ny, Nx,
               L.ptr(a_out), st)
    with Timer() as tm:
        for _ in range(nlaunch):
            L.call("igm_conv3x3_fwd_bf16", L.ptr(a_in), L.ptr(cnn.fwd[0]), L.ptr(cnn.bias(1)), None, 1, Ny, Nx,
                   L.ptr(a_out), st)
    conv_us = tm.ms * 1e3 / nlaunch
    conv_flop = 2.0 * 9 * 32 * 32 * Ny * Nx
    conv_tflops = conv_flop / (conv_us * 1e-6) / 1e12
    # energy + adjoint (mode 1) and the fused solver iteration (mode 2), each
    # kernel alone, back to back on the launch stream
    gU, gV = torch.empty_like(z), torch.empty_like(z)
    Uw, Vw = Ur.float().contiguous(), Vr.float().contiguous()
    for _ in range(5):
        op32.run_pass(1, Uw, Vw, gU=gU, gV=gV)
    with Timer() as tm:
        for _ in range(nlaunch):
            op32.run_pass(1, Uw, Vw, gU=gU, gV=gV)
    en_us = tm.ms * 1e3 / nlaunch
    en_gbs = BYTES_ENERGY_ADJOINT * Ny * Nx / (en_us * 1e-6) / 1e9
    mom = torch.zeros((nz, Ny, Nx, 4), dtype=torch.float32, device=dev)
    Uo, Vo = torch.empty_like(z), torch.empty_like(z)
    adam = L.IgmAdam(alpha=1e-3, beta1=0.9, beta2=0.999, epsilon=1e-7)
    for _ in range(5):
        op32.run_pass(2, Uw, Vw, Uo=Uo, Vo=Vo, moments=mom, adam=adam)
    with Timer() as tm:
        for _ in range(nlaunch):
            op32.run_pass(2, Uw, Vw, Uo=Uo, Vo=Vo, moments=mom, adam=adam)
    fused_us = tm.ms * 1e3 / nlaunch
    fused_gbs = BYTES_FUSED_ITERATION * Ny * Nx / (fused_us * 1e-6) / 1e9
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
    log(f"[bench] conv3x3 {conv_us:.1f} us/launch = {conv_tflops:.1f} TFLOP/s; energy+adjoint {en_us:.1f} us = "
        f"{en_gbs:.0f} GB/s; fused Adam pass {fused_us:.1f} us = {fused_gbs:.0f} GB/s; solver iteration "
        f"(pass + finalize + gaps) {it_ms * 1e3:.1f} us; device copy {copy_gbs:.0f} GB/s")

    # the dominant kernel of the timed step: the fused Adam energy pass when
    # the field is polished, the 3x3 conv otherwise
    if n_ref > 0:
        roofline = {"bound": "hbm", "achieved": round(fused_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": round(fused_gbs / PEAK_HBM_GBS, 4), "traffic": None,
                    "kernel": "energy_pass_kernel<float, 10, 2, 0> (fused energy + adjoint + Adam, 1000x1000)",
                    "bytes_per_cell": BYTES_FUSED_ITERATION, "us_per_launch": round(fused_us, 2),
                    "launches_per_step": int(n_ref)}
    else:
        roofline = {"bound": "mfma", "achieved": round(conv_tflops, 2), "peak": PEAK_BF16_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(conv_tflops / PEAK_BF16_TFLOPS, 4), "traffic": None,
                    "kernel": "conv3x3_mfma_kernel<0> (32->32, 1000x1000)", "us_per_launch": round(conv_us, 2),
                    "launches_per_step": 15}

    # ----------------------------------------------------- CPU baseline --
    cpu = None
    if not args.no_cpu:
        from oracle import cnn as OC
        ncore = os.cpu_count() or 1
        torch.set_num_threads(min(16, ncore))
        X = torch.stack([torch.as_tensor(g[k]) for k in keys], -1)[None]
        weights = cnn.get_weights()
        OC.forward(X[:, :64, :64], weights)
        tc = time.time()
        with torch.no_grad():
            OC.forward(X, weights)
        cpu_inf_s = time.time() - tc
        cpu = {"value": 1.0 / cpu_inf_s, "unit": "velocity fields/s", "cores": torch.get_num_threads(),
               "kind": "port",
               "sample": "one full 1000x1000 emulator inference (17 conv layers, fp32) of the oracle CNN "
                         "(torch-CPU restatement of emulate/neural_network.py); TensorFlow is not installed"}
        log(f"[bench] CPU oracle CNN inference {cpu_inf_s:.2f}s on {torch.get_num_threads()} threads")

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
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": "domes1000: 1000x1000 grid, dx=50 m, Nz=10, 20 seeded domes on a random-field bed; "
                               "emulated iceflow update (bf16 MFMA CNN 16x3x3x32 + 1x1, fp32 first layer) "
                               "polished by the shortest warm-started fp32 energy-solver schedule that reaches "
                               "<= 1 % rel err vs the float64 minimiser, plus the float64 energy-gap check",
                   "weights": os.path.basename(wpath),
                   "polish_its": int(n_ref), "polish_lr0": lr_ref,
                   "fallback_check_its": int(ice.emulator.fallback.check_its)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "emulated_rel_err_vs_fp64": round(emu_err, 5),
        "emulator_inference": {"ms": round(inf_ms, 4), "fields_per_s": round(1000.0 / inf_ms, 1),
                               "rel_err_vs_fp64": round(inf_err, 5)},
        "fallbacks": int(state.fallback_count),
        "solver_fields_per_s": round(solver_fps, 4),
        "solver_schedule": {"lr0": lr1, "flat": int(n_flat1), "decay_to_1e-3": int(n1 - n_flat1),
                            "rel_err": round(solver_err, 5)},
        "solver_iters_to_1pct": int(n1),
        "solver_reached_1pct": bool(solver_reached),
        "solver_reference_settings": ref_settings,
        "solver_us_per_iter": round(it_ms * 1e3, 2),
        "emulator_vs_solver_speedup": round(emu_fps / solver_fps, 1),
        "solver_early_stopping": {"iterations": int(res_es.iterations), "rel_err": round(es_err, 5)},
        "roofline_energy_adjoint": {"bound": "hbm", "achieved": round(en_gbs, 1), "peak": PEAK_HBM_GBS,
                                    "unit": "GB/s", "frac": round(en_gbs / PEAK_HBM_GBS, 4),
                                    "bytes_per_cell": BYTES_ENERGY_ADJOINT, "us_per_launch": round(en_us, 2)},
        "roofline_conv3x3": {"bound": "mfma", "achieved": round(conv_tflops, 2), "peak": PEAK_BF16_TFLOPS,
                             "unit": "TFLOP/s", "frac": round(conv_tflops / PEAK_BF16_TFLOPS, 4),
                             "us_per_launch": round(conv_us, 2),
                             "hbm_gbs": round(2 * 64.0 * Ny * Nx / (conv_us * 1e-6) / 1e9, 1)},
        "solver_iteration_us": round(it_ms * 1e3, 2),
        "device_copy_gbs": round(copy_gbs, 1),
        "fp64_reference": {"J": J64, "its": args.ref_flat + args.ref_decay, "change_under_refinement": ref_change},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
