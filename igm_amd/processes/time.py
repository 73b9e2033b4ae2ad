"""Drop-in ``time`` process (igm/processes/time/time.py:11-65): CFL-adaptive
time step from the GPU max-|velbar| reduction; scalar logic on the host."""
import numpy as np
import torch

from .. import _lib as L


def initialize(cfg, state):
    """time.py:11-26.  t, dt and the save times are float32, as the
    reference's tf.Variable / tf.constant(float32): the save-time snap
    t + (t_save - t) then lands exactly on t_save (Sterbenz)."""
    c = cfg.processes.time
    f = np.float32
    state.t = f(c.start)
    state.itsave = -1
    state.dt = f(c.step_max)
    state.dt_target = f(c.step_max)
    state.time_save = np.array(np.ndarray.tolist(np.arange(c.start, c.end, c.save)) + [c.end], np.float32)
    state._absmax_ws = torch.empty(1024, dtype=torch.float32, device=state.thk.device)
    state._absmax = torch.empty(1, dtype=torch.float32, device=state.thk.device)


def velomax(state):
    n = state.ubar.numel()
    L.call("igm_absmax2_f32", L.ptr(state.ubar.contiguous()), L.ptr(state.vbar.contiguous()), n,
           L.ptr(state._absmax_ws), L.ptr(state._absmax), L.stream_ptr())
    return float(state._absmax.item())


def update(cfg, state):
    """time.py:28-65, in float32."""
    c = cfg.processes.time
    f = np.float32
    if hasattr(state, "logger"):
        state.logger.info("Update DT at time : " + str(state.t))
    if c.cfl > 0:
        vm = f(velomax(state))
        if vm > 0:
            state.dt_target = min(f(c.cfl) * f(state.dx) / vm, f(c.step_max))
        else:
            state.dt_target = f(c.step_max)
    else:
        state.dt_target = f(c.step_max)
    state.dt = state.dt_target
    if state.time_save[state.itsave + 1] <= f(state.t + state.dt):
        state.dt = f(state.time_save[state.itsave + 1] - state.t)
        state.saveresult = True
        state.itsave += 1
    else:
        state.saveresult = False
    if state.it >= 0:
        state.t = f(state.t + state.dt)
    state.continue_run = bool(state.t < f(c.end))


def finalize(cfg, state):
    pass
