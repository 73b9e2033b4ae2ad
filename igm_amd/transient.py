"""Transient glacier evolution: the runner's time loop over the GPU processes.

IGM's runner calls each process's update in the experiment's order inside
``while state.continue_run`` (igm/common/runner/modules/src.py:18-41); the
transient experiments order them smb_simple, iceflow, time, thk
(tests/test_full_glacier_evolution_oggm_shop/experiment/params.yaml:7-13).
``run`` does the same with the drop-in modules of this package and records
the ice volume, the time step and the iceflow wall time per step.
"""
import time as _time

import numpy as np
import torch

from . import config, iceflow
from .fields import as_device
from .processes import smb_simple, thk as thkmod, time as timemod


class State:
    pass


def make_state(fields):
    """State from a synthetic field dict (thk, topg, usurf, arrhenius, slidingco, dX, dx)."""
    st = State()
    for k in ("thk", "topg", "arrhenius", "slidingco", "dX"):
        setattr(st, k, as_device(fields[k]))
    st.usurf = as_device(fields["usurf"])
    st.dx = float(fields["dx"])
    st.it = -1
    return st


def run(cfg, state, log=None):
    """Run until cfg.processes.time.end; returns a dict of per-step records."""
    modules = [smb_simple, iceflow, timemod, thkmod]
    thkmod.initialize(cfg, state)
    timemod.initialize(cfg, state)
    smb_simple.initialize(cfg, state)
    iceflow.initialize(cfg, state)
    rec = dict(t=[], dt=[], volume=[], iceflow_ms=[], fallbacks=[])
    state.continue_run = True
    state.it = 0
    dx2 = float(state.dx) ** 2
    while state.continue_run:
        for m in modules:
            if m is iceflow:
                torch.cuda.synchronize()
                t0 = _time.perf_counter()
                m.update(cfg, state)
                torch.cuda.synchronize()
                rec["iceflow_ms"].append((_time.perf_counter() - t0) * 1e3)
            else:
                m.update(cfg, state)
        rec["t"].append(float(state.t))
        rec["dt"].append(float(state.dt))
        rec["volume"].append(float(torch.sum(state.thk, dtype=torch.float64)) * dx2 / 1e9)
        rec["fallbacks"].append(int(getattr(state, "fallback_count", 0)))
        if log is not None and state.it % 10 == 0:
            log(f"it {state.it} t {state.t:.2f} dt {state.dt:.3f} volume {rec['volume'][-1]:.4f} km3 "
                f"iceflow {rec['iceflow_ms'][-1]:.2f} ms")
        state.it += 1
    return {k: np.asarray(v) for k, v in rec.items()}


def transient_cfg(method="emulated", years=200.0, retrain_freq=10, check_its=50, check_freq=10):
    """The transient workload's settings: CFL 0.3, step_max 1 yr, SMB =
    min(0.005 (usurf - 2800), 2) m/yr, online retraining every 10 steps, and
    the float64 energy-gap check every ``check_freq`` emulated updates
    (``check_its`` solver iterations from the emulated field give the
    comparison energy; a gap above 1 % falls back to a solve)."""
    cfg = config.default_cfg()
    cfg.processes.iceflow.method = method
    cfg.processes.iceflow.emulator.retrain_freq = retrain_freq
    fb = cfg.processes.iceflow.emulator.fallback
    fb.check_its = int(check_its)
    fb.check_freq = int(check_freq)
    cfg.processes.time.start = 0.0
    cfg.processes.time.end = float(years)
    cfg.processes.time.save = float(years)
    cfg.processes.time.cfl = 0.3
    cfg.processes.time.step_max = 1.0
    cfg.processes.smb_simple.array = [["time", "gradabl", "gradacc", "ela", "accmax"],
                                      [0.0, 0.005, 0.005, 2800.0, 2.0],
                                      [1.0e6, 0.005, 0.005, 2800.0, 2.0]]
    return cfg


# years at which the two runs' ice volumes are compared (linear interpolation
# between their own time steps)
REPORT_YEARS = (1, 2, 5, 10, 15, 20, 30, 50, 100, 200)


def main():
    """``python -m igm_amd.transient --years 200``: the 1500 x 1500 transient
    workload, emulated (online retraining + fallback check) vs solved; prints
    one JSON line with the final volumes and the iceflow fields/s of each."""
    import argparse
    import json
    import os

    from . import emulator as EM
    from . import synthetic as syn
    ap = argparse.ArgumentParser()
    ap.add_argument("--years", type=float, default=200.0)
    ap.add_argument("--size", type=int, default=1500)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--methods", nargs="*", default=["emulated", "solved"])
    ap.add_argument("--check-its", type=int, default=50)
    ap.add_argument("--check-freq", type=int, default=10)
    args = ap.parse_args()
    g = syn.domes1500() if args.size == 1500 else syn.domes1000()
    out = {"workload": f"domes{args.size}", "years": args.years}
    for method in args.methods:
        cfg = transient_cfg(method=method, years=args.years, check_its=args.check_its,
                            check_freq=args.check_freq)
        if args.weights:
            cfg.processes.iceflow.emulator.name = os.path.abspath(args.weights)
        st = make_state(g)
        t0 = _time.perf_counter()
        rec = run(cfg, st, log=lambda s: print(f"[{method}] {s}", flush=True))
        wall = _time.perf_counter() - t0
        out[method] = {"steps": int(len(rec["t"])), "final_volume_km3": float(rec["volume"][-1]),
                       "initial_volume_km3": float(rec["volume"][0]),
                       "iceflow_fields_per_s": float(1000.0 / np.mean(rec["iceflow_ms"])),
                       "iceflow_ms_median": float(np.median(rec["iceflow_ms"])),
                       "wall_s": wall, "fallbacks": int(rec["fallbacks"][-1]),
                       "volume_at": {f"{t:g}": float(np.interp(t, rec["t"], rec["volume"]))
                                     for t in REPORT_YEARS if t <= args.years}}
    if "emulated" in out and "solved" in out:
        ve, vs = out["emulated"]["volume_at"], out["solved"]["volume_at"]
        out["volume_rel_diff_at"] = {t: (abs(ve[t] - vs[t]) / vs[t] if vs[t] > 0 else None) for t in vs}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
