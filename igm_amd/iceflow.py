"""Drop-in ``iceflow`` process module: initialize / update / finalize.

Same contract as igm/processes/iceflow/iceflow.py:46-119 (validated by
igm/common/runner/modules/validator.py:1-9): reads ``state.thk``,
``state.usurf``, ``state.dX`` (+ ``arrhenius``, ``slidingco``, ``U``, ``V``
when present) and rebinds ``state.U``, ``state.V`` (Nz, Ny, Nx) and
``uvelbase, vvelbase, uvelsurf, vvelsurf, ubar, vbar`` (Ny, Nx), plus
``vert_weight``, ``levels``, ``zeta``, ``dzeta``, ``PAD`` and the cost
records.  Fields are torch tensors on the GPU (numpy inputs are accepted and
moved there).  All field arithmetic runs in the HIP kernels of
libigm_amd.so; this module is host control flow.

Methods (cfg.processes.iceflow.method):
* ``solved``     -- the plain energy solver (solve/solve.py:169-188);
* ``emulated``   -- the CNN emulator with periodic physics-informed
  retraining (emulate/emulate.py:116-281), plus an optional warm-started
  solver polish of the emulated field (cfg.processes.iceflow.emulator.refine)
  and an energy-gap fallback to the solver
  (cfg.processes.iceflow.emulator.fallback), both absent from the reference;
* ``diagnostic`` -- emulated, plus a periodic warm-started solve that logs
  the emulator-vs-solver misfit (diagnostic/diagnostic.py:16-84).
"""
import os

import numpy as np
import torch

from . import config as cfgmod
from . import emulator as emu
from . import solver as slv
from .energy import EnergyOperator, physics_struct
from .fields import as_device, velocity_outputs
from .vert_disc import compute_levels, compute_zeta_dzeta, define_vertical_weight

FIELDIN = ("thk", "usurf", "arrhenius", "slidingco", "dX")


def _check_supported(cfg):
    ice = cfg.processes.iceflow
    if ice.numerics.vert_basis != "Lagrange":
        if ice.numerics.vert_basis in ("Legendre", "SIA"):
            raise NotImplementedError(f"vert_basis {ice.numerics.vert_basis!r} is not supported")
        raise ValueError(f"Unknown vertical basis: {ice.numerics.vert_basis}")  # iceflow.py:82
    if int(ice.numerics.staggered_grid) != 1:
        raise NotImplementedError("only staggered_grid=1 is supported")
    if list(ice.physics.energy_components) != ["shear", "sliding_weertman", "gravity"]:
        raise NotImplementedError("energy_components must be ['shear', 'sliding_weertman', 'gravity']")
    if len(ice.physics.sliding_law) > 0:
        raise NotImplementedError("sliding_law is not supported")
    if int(ice.physics.dim_arrhenius) != 2:
        raise NotImplementedError("dim_arrhenius must be 2")
    if ice.physics.force_negative_gravitational_energy:
        raise NotImplementedError("force_negative_gravitational_energy is not supported")
    if ice.solver.lbfgs:
        # solve.py:171-172
        raise ValueError("solve_iceflow_lbfgs formely implemented, not working yet, will be updated.")
    if ice.emulator.exclude_borders != 0 or ice.emulator.network.multiple_window_size != 0:
        raise NotImplementedError("exclude_borders / multiple_window_size are not supported")
    if ice.emulator.network.architecture != "cnn" or ice.emulator.network.cnn3d_for_vertical:
        raise NotImplementedError("only the 'cnn' emulator architecture is supported")


def _fields_to_device(state):
    for f in FIELDIN + ("U", "V"):
        if hasattr(state, f):
            setattr(state, f, as_device(getattr(state, f)))


def initialize_iceflow_fields(cfg, state):
    """iceflow/utils.py:15-32."""
    ph = cfg.processes.iceflow.physics
    Nz = int(cfg.processes.iceflow.numerics.Nz)
    if not hasattr(state, "arrhenius"):
        state.arrhenius = torch.ones_like(state.thk) * (ph.init_arrhenius * ph.enhancement_factor)
    if not hasattr(state, "slidingco"):
        state.slidingco = torch.ones_like(state.thk) * ph.init_slidingco
    if not hasattr(state, "U"):
        state.U = torch.zeros((Nz,) + tuple(state.thk.shape), dtype=torch.float32, device=state.thk.device)
        state.V = torch.zeros_like(state.U)


def initialize_iceflow_solver(cfg, state):
    """solve.py:11-20: one optimizer for all solves (its step counter persists)."""
    state.optimizer = slv.LegacyAdamState(lr=float(cfg.processes.iceflow.solver.step_size))


def energy_operator(cfg, state, dtype=torch.float32):
    """The energy operator of the state's fields.  The float32 one is kept on
    the state and reused while its inputs are unchanged: the same field
    tensors (address and version -- torch counts every in-place write), the
    same vertical grid and physics.  The cache entry holds the field tensors,
    so no other tensor can take their memory, and address, meanwhile."""
    fields = {f: getattr(state, f) for f in FIELDIN}
    if dtype != torch.float32:
        fields = {k: v.to(dtype) for k, v in fields.items()}
        return EnergyOperator(np.asarray(state.dzeta), cfg.processes.iceflow.physics,
                              fields=fields, dtype=dtype)
    dz = np.asarray(state.dzeta, dtype=np.float64)
    phys = cfg.processes.iceflow.physics
    key = (tuple((f.data_ptr(), f._version, tuple(f.shape), f.dtype, f.device)
                 for f in fields.values()),
           dz.tobytes(), bytes(physics_struct(phys)))
    cached = getattr(state, "_energy_op", None)
    if cached is not None and cached[0] == key:
        return cached[1]
    op = EnergyOperator(dz, phys, fields=fields, dtype=dtype)
    state._energy_op = (key, op, tuple(fields.values()))
    return op


def solve_iceflow(cfg, state, U, V, nbitmax=None):
    """solve.py:22-119 on the GPU.  Returns (U, V) unmasked and the cost list."""
    ice = cfg.processes.iceflow
    op = energy_operator(cfg, state)
    res = slv.solve(op, U, V, int(ice.solver.nbitmax if nbitmax is None else nbitmax), state.optimizer,
                    early_stopping=True, check_every=int(cfgmod.get(ice, "solver.check_every", 16)))
    return res


def _set_outputs(cfg, state, out):
    for k, v in out.items():
        setattr(state, k, v)


PUBLISHED = ("U", "V", "uvelbase", "vvelbase", "uvelsurf", "vvelsurf", "ubar", "vbar")


def publish_outputs(cfg, state):
    """Hand the output fields to the rest of the run in the form
    ``cfg.processes.iceflow.output_format`` names: "torch" (default: device
    tensors, for GPU consumers such as this package's thk, time and
    vert_flow processes) or "numpy" (host float32 arrays: IGM's own
    TensorFlow processes take them as they are, e.g. tf.abs(state.ubar) in
    time.py:36-39 and the flux divergence in thk.py:36-38).  Inputs given
    back as numpy at the next update are moved to the GPU again."""
    fmt = cfgmod.get(cfg.processes.iceflow, "output_format", "torch")
    if fmt == "torch":
        return
    if fmt != "numpy":
        raise ValueError(f"output_format must be 'torch' or 'numpy', not {fmt!r}")
    for k in PUBLISHED:
        v = getattr(state, k, None)
        if isinstance(v, torch.Tensor):
            setattr(state, k, v.detach().to("cpu", torch.float32).numpy())


def update_iceflow_solved(cfg, state):
    """solve.py:169-188."""
    ice = cfg.processes.iceflow
    res = solve_iceflow(cfg, state, state.U, state.V)
    out = velocity_outputs(res.U, res.V, state.thk, state.vert_weight, int(ice.numerics.Nz))
    _set_outputs(cfg, state, out)
    if ice.force_max_velbar > 0:
        raise NotImplementedError("force_max_velbar is not supported")
    if len(ice.solver.save_cost) > 0:
        np.savetxt(ice.emulator.output_directory + ice.solver.save_cost + "-" + str(state.it) + ".dat",
                   np.asarray(res.costs), fmt="%5.10f")
    state.COST_Glen = float(res.costs[-1]) if len(res.costs) else float("nan")
    state.Cost_Glen = res.costs
    state.solver_iterations = res.iterations


def initialize(cfg, state):
    """iceflow.py:46-95."""
    if hasattr(state, "was_initialize_iceflow_already_called"):
        return
    _check_supported(cfg)
    _fields_to_device(state)
    initialize_iceflow_fields(cfg, state)
    ice = cfg.processes.iceflow
    if ice.method == "emulated":
        emu.initialize_iceflow_emulator(cfg, state)
    elif ice.method == "solved":
        initialize_iceflow_solver(cfg, state)
    elif ice.method == "diagnostic":
        emu.initialize_iceflow_emulator(cfg, state)
        initialize_iceflow_solver(cfg, state)
        state.diagno = []
        state.tlast_diagno = -(10 ** 10)
    else:
        raise ValueError(f"Unknown iceflow method: {ice.method}")
    if ice.method == "emulated" and cfgmod.get(ice, "emulator.fallback.enabled", False):
        initialize_iceflow_solver(cfg, state)

    Nz, vs = int(ice.numerics.Nz), float(ice.numerics.vert_spacing)
    state.vert_weight = define_vertical_weight(Nz, vs)
    state.levels = compute_levels(Nz, vs)
    state.zeta, state.dzeta = compute_zeta_dzeta(state.levels)
    state.Leg_P, state.Leg_dPdz, state.Leg_I = None, None, None
    state.PAD = [[0, 0], [0, 0], [0, 0], [0, 0]]

    if not ice.method == "solved":
        emu.update_iceflow_emulator(cfg, state, 0)
        emu.update_iceflow_emulated(cfg, state)
    publish_outputs(cfg, state)
    state.was_initialize_iceflow_already_called = True


def update(cfg, state):
    """iceflow.py:97-113 (+ the energy-gap fallback for 'emulated')."""
    if hasattr(state, "logger"):
        state.logger.info("Update ICEFLOW at iteration : " + str(state.it))
    _fields_to_device(state)
    ice = cfg.processes.iceflow
    if ice.method in ("emulated", "diagnostic"):
        if (ice.emulator.retrain_freq > 0) & (state.it > 0):
            emu.update_iceflow_emulator(cfg, state, state.it, pertubate=ice.emulator.pertubate)
        emu.update_iceflow_emulated(cfg, state)
        if ice.method == "emulated" and int(cfgmod.get(ice, "emulator.refine.its", 0)) > 0:
            refine_emulated(cfg, state)
        if ice.method == "emulated" and cfgmod.get(ice, "emulator.fallback.enabled", False):
            fallback_check(cfg, state)
        if ice.method == "diagnostic":
            update_iceflow_diagnostic(cfg, state)
    elif ice.method == "solved":
        update_iceflow_solved(cfg, state)
    publish_outputs(cfg, state)


def finalize(cfg, state):
    """iceflow.py:116-119."""
    if cfg.processes.iceflow.emulator.save_model:
        emu.save_iceflow_model(cfg, state)


# ------------------------------------------------------------- refine --

def refine_emulated(cfg, state):
    """Polish the emulated field: ``refine.its`` plain-solver iterations
    warm-started from state.U, state.V with a fresh legacy-Adam state: a
    geometric warm-up from ``warmup_lr`` to ``refine.lr`` over the first
    ``warmup_frac`` of them (none by default), step ``refine.lr`` for the next
    ``flat_frac``, then a geometric decay to ``lr_end``, with Adam's
    ``beta1``, ``beta2`` and ``epsilon``; the outputs are
    rebuilt from the polished field.  A network whose error is a few percent
    reaches the solver's accuracy in a small fraction of the iterations a
    solve from zero needs."""
    ice = cfg.processes.iceflow
    d = cfgmod.ICEFLOW["emulator"]["refine"]
    rf = {k: cfgmod.get(ice, "emulator.refine." + k, v) for k, v in d.items()}
    n = int(rf["its"])
    if n <= 0:
        return
    op = energy_operator(cfg, state)
    lr = float(rf["lr"])
    n_warm = int(float(rf["warmup_frac"]) * n)
    n_flat = int(float(rf["flat_frac"]) * n)
    lr_start = float(rf["warmup_lr"]) if float(rf["warmup_lr"]) > 0 else lr
    opt = slv.LegacyAdamState(lr=lr, beta1=float(rf["beta1"]), beta2=float(rf["beta2"]),
                              epsilon=float(rf["epsilon"]))
    res = slv.solve(op, state.U, state.V, n, opt, early_stopping=False,
                    lr_schedule=slv.adam_schedule(lr, n_flat, n - n_warm - n_flat, float(rf["lr_end"]),
                                                  n_warmup=n_warm, lr_start=lr_start))
    out = velocity_outputs(res.U, res.V, state.thk, state.vert_weight, int(ice.numerics.Nz))
    _set_outputs(cfg, state, out)
    state.iceflow_unmasked = (res.U, res.V)
    state.COST_REFINED = res.costs_device[-1]  # float64 0-dim device tensor (no host wait)


# ----------------------------------------------------------- fallback --

def _near_ice(state, dtype):
    """Nodes with ice in one of their cells: the 3 x 3 neighbourhood maximum of
    thk > 0.  Kept on the state while thk is the same tensor at the same
    version (the entry holds the tensor, so its address stays its own)."""
    thk = state.thk
    key = (thk.data_ptr(), thk._version, tuple(thk.shape), dtype)
    cached = getattr(state, "_near_ice", None)
    if cached is not None and cached[0] == key:
        return cached[1]
    ice = (thk > 0).to(dtype)[None, None]
    near = torch.nn.functional.max_pool2d(ice, 3, stride=1, padding=1)[0, 0] > 0
    state._near_ice = (key, near, thk)
    return near


def fallback_check(cfg, state):
    """Energy-gap fallback of the emulated field to the plain solver.

    Every ``check_freq`` calls: J_emu = J(U, V) of the emulated field before
    the thk > 0 mask of the outputs (emulate.py:143-144) -- the network's own
    output, or the polished field -- float64 reduction: the energy the
    network is trained on and the solver minimises.  (The mask zeroes the
    ice-free nodes of the margin cells and adds an artificial margin shear
    that a few solver steps remove, so masked energies are not compared.)
    Nodes with no ice within one node (all their cells ice-free) are set to
    zero first: they enter only the sliding term of ice-free cells, which is
    smallest at zero, so this is their optimum and no ice cell sees them,
    while left as they are (the network's output there is unconstrained,
    Adam leaves steps of its size there) they would dominate the gap.
    J_check is the lowest COST seen by ``check_its`` solver iterations
    started from that field.  If J_emu > J_check + rel_energy_gap *
    |J_check|, the solve continues (up to ``max_its`` iterations with early
    stopping) and its field, masked, replaces the emulated one.  With
    check_its = 0 the check only records J_emu.  Records
    state.COST_EMULATED (a float64 0-dim device tensor), state.COST_CHECK
    (float) and state.fallback_count.
    """
    fb = cfg.processes.iceflow.emulator.fallback
    state.fallback_calls = getattr(state, "fallback_calls", 0) + 1
    state.fallback_count = getattr(state, "fallback_count", 0)
    if (state.fallback_calls - 1) % max(1, int(fb.check_freq)) != 0:
        return
    op = energy_operator(cfg, state)
    full = getattr(state, "iceflow_unmasked", None)
    if full is None:
        Y = getattr(state, "iceflow_Y", None)
        if Y is None:
            full = (state.U, state.V)
        else:
            nz = int(cfg.processes.iceflow.numerics.Nz)
            full = (Y[0, ..., :nz].permute(2, 0, 1).contiguous(), Y[0, ..., nz:].permute(2, 0, 1).contiguous())
    near_ice = _near_ice(state, full[0].dtype)
    zero = torch.zeros((), dtype=full[0].dtype, device=full[0].device)
    U0 = torch.where(near_ice, full[0], zero).contiguous()
    V0 = torch.where(near_ice, full[1], zero).contiguous()
    k = int(fb.check_its)
    if k <= 0:
        # nothing to decide: record J_emu (the host does not wait for the GPU)
        state.COST_EMULATED = op.energy(U0, V0)[0, 3]  # float64 0-dim device tensor
        return
    res = slv.solve(op, U0, V0, k, state.optimizer, early_stopping=False)
    # the COST of the first iteration is the energy of the field it starts
    # from: J_emu (float64 0-dim device tensor); one read of the history
    state.COST_EMULATED = res.costs_device[0]
    costs = res.costs
    j_emu = float(costs[0])
    j_chk = float(min(np.min(costs), j_emu))
    state.COST_CHECK = j_chk
    if j_emu > j_chk + float(fb.rel_energy_gap) * abs(j_chk):
        state.fallback_count += 1
        res2 = slv.solve(op, res.U, res.V, int(fb.max_its), state.optimizer, early_stopping=True)
        out = velocity_outputs(res2.U, res2.V, state.thk, state.vert_weight,
                               int(cfg.processes.iceflow.numerics.Nz))
        _set_outputs(cfg, state, out)
        state.COST_Glen = float(res2.costs[-1]) if len(res2.costs) else j_chk


# --------------------------------------------------------- diagnostic --

def update_iceflow_diagnostic(cfg, state):
    """diagnostic/diagnostic.py:25-71: every 10 years, solve warm-started
    from the emulated field and log the misfit and both costs."""
    t = float(getattr(state, "t", 0.0))
    if (t - state.tlast_diagno) < 10:
        return
    res = solve_iceflow(cfg, state, state.U, state.V)
    Nz = int(cfg.processes.iceflow.numerics.Nz)
    sol = velocity_outputs(res.U, res.V, state.thk, state.vert_weight, Nz)
    l1, l2 = computemisfit(state.thk, state.ubar - sol["ubar"], state.vbar - sol["vbar"])
    vol = float(torch.sum(state.thk, dtype=torch.float64)) * float(state.dX[0, 0]) ** 2 / 1e9
    cost_emu = float(state.COST_EMULATOR[-1]) if getattr(state, "COST_EMULATOR", None) else float("nan")
    state.diagno.append([t, l1, l2, float(res.costs[-1]), cost_emu, res.iterations,
                         len(getattr(state, "COST_EMULATOR", [])), vol])
    np.savetxt("errors_diagno.txt", np.stack(state.diagno), delimiter=",", fmt="%10.3f",
               header="time,l1,l2,COST_Glen,COST_Emulator,nb_it_solve,nb_it_emul,vol", comments="")
    state.tlast_diagno = t


def computemisfit(thk, dubar, dvbar):
    """diagnostic.py:74-84 on the velbar difference (masked thk > 1)."""
    m = (thk > 1).to(torch.float64)
    n = 2.0 * float(torch.sum(m))
    if n == 0:
        return 0.0, 0.0
    a = torch.abs(dubar.double()) * m
    b = torch.abs(dvbar.double()) * m
    l1 = float(torch.sum(a) + torch.sum(b)) / n
    l2 = float(torch.sum(a * a) + torch.sum(b * b)) / n
    return l1, float(np.sqrt(l2))
