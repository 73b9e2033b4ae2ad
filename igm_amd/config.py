"""Default configuration tree for the ice-flow process and its siblings.

Mirrors the keys and defaults of the reference's
igm/conf/processes/iceflow.yaml:1-72, thk.yaml, time.yaml and
smb_simple.yaml, as an attribute-access tree (any object with the same
attributes -- an OmegaConf DictConfig for instance -- works the same).

Keys this build adds (absent from the reference) live under
``processes.iceflow.emulator.fallback``, ``processes.iceflow.emulator.refine``
and ``processes.iceflow.solver``
(``check_every``); each is documented where it is read.
"""
import copy
from types import SimpleNamespace

ICEFLOW = {
    "method": "emulated",
    "force_max_velbar": 0.0,
    # build-specific: "torch" (device tensors) or "numpy" (host arrays for
    # TensorFlow consumers of the outputs)
    "output_format": "torch",
    "physics": {
        "energy_components": ["shear", "sliding_weertman", "gravity"],
        "sliding_law": "",
        "gravity_cst": 9.81,
        "ice_density": 910.0,
        "init_slidingco": 0.0464,
        "init_arrhenius": 78.0,
        "enhancement_factor": 1.0,
        "exp_glen": 3.0,
        "exp_weertman": 3.0,
        "regu_glen": 1.0e-05,
        "regu_weertman": 1.0e-10,
        "dim_arrhenius": 2,
        "thr_ice_thk": 0.1,
        "min_sr": 1.0e-20,
        "max_sr": 1.0e20,
        "force_negative_gravitational_energy": False,
        "cf_eswn": [],
    },
    "numerics": {
        "Nz": 10,
        "vert_spacing": 4.0,
        "staggered_grid": 1,
        "vert_basis": "Lagrange",
    },
    "solver": {
        "step_size": 1.0,
        "nbitmax": 100,
        "stop_if_no_decrease": True,
        "optimizer": "Adam",
        "lbfgs": False,
        "save_cost": "",
        "plot_sol": False,
        # build-specific: the host reads the device early-stopping flag every
        # check_every iterations (the result does not depend on it)
        "check_every": 16,
    },
    "emulator": {
        "fieldin": ["thk", "usurf", "arrhenius", "slidingco", "dX"],
        "retrain_freq": 10,
        "lr": 2.0e-05,
        "lr_init": 0.0001,
        "lr_decay": 0.95,
        "warm_up_it": -10000000000.0,
        "nbit_init": 1,
        "nbit": 1,
        "framesizemax": 750,
        "split_patch_method": "sequential",
        "pretrained": True,
        "name": "",
        "save_model": False,
        "exclude_borders": 0,
        "optimizer": "Adam",
        "optimizer_clipnorm": 1.0,
        "optimizer_epsilon": 1.0e-07,
        "save_cost": "",
        "output_directory": "",
        "plot_sol": False,
        "pertubate": False,
        "network": {
            "architecture": "cnn",
            "multiple_window_size": 0,
            "activation": "LeakyReLU",
            "nb_layers": 16,
            "nb_blocks": 4,
            "nb_out_filter": 32,
            "conv_ker_size": 3,
            "dropout_rate": 0,
            "weight_initialization": "glorot_uniform",
            "cnn3d_for_vertical": False,
        },
        # build-specific energy-gap fallback to the plain solver: every
        # check_freq update() calls, check_its solver iterations (the
        # solver's own step_size) from the emulated field (before the
        # thk > 0 output mask) give J_check, the lowest energy they see; the
        # solve continues up to max_its iterations (with EarlyStopping) and
        # replaces the emulated field when
        # J_emulated > J_check + rel_energy_gap * |J_check|
        "fallback": {
            "enabled": True,
            "rel_energy_gap": 0.01,
            "check_its": 10,
            "check_freq": 1,
            "max_its": 2000,
        },
        # build-specific polish of every emulated field: `its` plain-solver
        # iterations warm-started from it, step lr for flat_frac of them,
        # then a geometric decay to lr_end (0 its: off); beta1, beta2 and
        # epsilon of its Adam (the Keras defaults unless set)
        "refine": {"its": 0, "lr": 3.0, "flat_frac": 0.3, "lr_end": 1.0e-3, "beta1": 0.9,
                   "beta2": 0.999, "epsilon": 1.0e-7, "warmup_frac": 0.0, "warmup_lr": 0.0},
    },
}

THK = {"slope_type": "superbee", "ratio_density": 0.910, "default_sealevel": 0.0}
TIME = {"start": 2000.0, "end": 2100.0, "save": 10.0, "cfl": 0.3, "step_max": 1.0}
SMB_SIMPLE = {"update_freq": 1.0, "file": "param.txt", "array": []}


def _ns(d):
    if isinstance(d, dict):
        return SimpleNamespace(**{k: _ns(v) for k, v in d.items()})
    return copy.deepcopy(d)


def default_cfg():
    """A fresh cfg with cfg.processes.{iceflow, thk, time, smb_simple}."""
    return SimpleNamespace(processes=SimpleNamespace(
        iceflow=_ns(ICEFLOW), thk=_ns(THK), time=_ns(TIME), smb_simple=_ns(SMB_SIMPLE)))


def get(node, dotted, default=None):
    """Read an optional key (build-specific keys may be absent from a
    reference cfg)."""
    cur = node
    for part in dotted.split("."):
        if not hasattr(cur, part):
            return default
        cur = getattr(cur, part)
    return cur
