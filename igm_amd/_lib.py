"""ctypes binding of libigm_amd.so (the C-ABI declared in include/igm_amd.h).

The library is loaded from the package directory.  There is no fallback: if
it is missing or fails to load, every entry point raises.
"""
import ctypes as C
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libigm_amd.so")

MAX_NZ = 32
VP = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
F64 = C.c_double


class IgmPhysics(C.Structure):
    _fields_ = [
        ("exp_glen", F64), ("exp_weertman", F64), ("regu_glen", F64),
        ("regu_weertman", F64), ("thr_ice_thk", F64), ("min_sr", F64), ("max_sr", F64),
        ("ice_density", F64), ("gravity_cst", F64),
        ("force_negative_gravitational_energy", I32),
    ]


class IgmAdam(C.Structure):
    _fields_ = [("alpha", F64), ("beta1", F64), ("beta2", F64), ("epsilon", F64)]


class IgmEarlyStop(C.Structure):
    _fields_ = [
        ("best_loss", F64), ("relative_min_delta", F64), ("has_best", I32), ("wait", I32),
        ("patience", I32), ("stopped", I32), ("stop_iter", I32), ("iters_done", I32),
        ("passes", I32),
    ]


EARLYSTOP_STOPPED_OFFSET = IgmEarlyStop.stopped.offset

# name -> (restype, argtypes)
_SIGS = {
    "igm_last_error": (C.c_char_p, []),
    "igm_energy_workspace_doubles": (I64, [I32, I32, I32]),
    "igm_energy_coeffs_f32": (C.c_int, [VP, VP, VP, VP, VP, I64, I64, I64, VP, I32, I32, I32, VP, VP, VP, VP]),
    "igm_energy_coeffs_f64": (C.c_int, [VP, VP, VP, VP, VP, I64, I64, I64, VP, I32, I32, I32, VP, VP, VP, VP]),
    "igm_energy_pass_f32": (C.c_int, [I32, I32, I32, I32, I32, VP, VP, I64, I64, I64, I64, VP, VP, VP, VP,
                                      VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, I32, VP]),
    "igm_energy_pass_f64": (C.c_int, [I32, I32, I32, I32, I32, VP, VP, I64, I64, I64, I64, VP, VP, VP, VP,
                                      VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, I32, VP]),
    "igm_energy_finalize": (C.c_int, [VP, I32, I32, I32, VP, VP, VP, I32, VP]),
    "igm_energy_pass2_f32": (C.c_int, [I32, I32, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP,
                                       VP, VP, I32, VP, VP, VP, VP, I32, VP]),
    "igm_energy_finalize_pair": (C.c_int, [VP, I32, I32, VP, VP, VP, I32, VP]),
    "igm_energy_eval_f32_noalloc": (C.c_int, [I32, I32, I32, I32, I32, VP, VP, I64, I64, I64, I64, VP, VP, VP,
                                              VP, VP, VP, VP, VP]),
    "igm_energy_solve_pass_f32": (C.c_int, [I32, I32, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP,
                                            I32, VP, VP, VP, VP, VP, VP, I32, VP]),
    "igm_velocity_outputs_f32": (C.c_int, [I32, I32, I32, VP, VP, I64, I64, I64, VP, VP, VP, VP, VP, VP, VP,
                                           VP, VP, VP, VP]),
    "igm_velocity_outputs_f64": (C.c_int, [I32, I32, I32, VP, VP, I64, I64, I64, VP, VP, VP, VP, VP, VP, VP,
                                           VP, VP, VP, VP]),
    "igm_conv_in_fwd": (C.c_int, [VP, VP, VP, VP, VP, I64, I64, I64, VP, I32, I32, I32, VP, VP, VP, VP]),
    "igm_conv3x3_fwd_bf16": (C.c_int, [VP, VP, VP, VP, I32, I32, I32, VP, VP]),
    "igm_conv3x3_dgrad_bf16": (C.c_int, [VP, VP, VP, VP, I32, I32, I32, VP, VP]),
    "igm_conv1x1_head_fwd": (C.c_int, [VP, VP, VP, VP, I32, I32, I32, I32, VP, VP]),
    "igm_conv1x1_head_bwd": (C.c_int, [VP, VP, VP, VP, I32, I32, I32, I32, VP, VP, VP]),
    "igm_wgrad_workspace_floats": (I64, [I32, I32]),
    "igm_conv_wgrad_bf16": (C.c_int, [I32, VP, VP, VP, I32, I32, I32, I32, I32, VP, I32, VP, VP, VP]),
    "igm_conv_in_wgrad": (C.c_int, [VP, VP, VP, VP, VP, I64, I64, I64, VP, I32, I32, I32, VP, I32, VP, I32,
                                    VP, VP, VP]),
    "igm_conv_wgrad_partials_bf16": (C.c_int, [I32, VP, VP, VP, I32, I32, I32, I32, VP, I32, VP]),
    "igm_conv_in_wgrad_partials": (C.c_int, [VP, VP, VP, VP, VP, I64, I64, I64, VP, I32, I32, I32, VP,
                                             I32, VP, I32, VP]),
    "igm_conv_wgrad_reduce_f32": (C.c_int, [I32, I32, I32, VP, I32, VP, VP, VP]),
    "igm_conv_in_dgrad": (C.c_int, [VP, VP, VP, I32, I32, I32, VP, VP]),
    "igm_vert_flow_kinematic_f32": (C.c_int, [I32, I32, I32, VP, VP, VP, VP, VP, VP, C.c_double, VP, VP, VP,
                                              VP]),
    "igm_adam_clip_f32": (C.c_int, [VP, VP, VP, VP, I64, VP, I32, VP, F64, VP, VP]),
    "igm_pack_conv_weights": (C.c_int, [VP, VP, I32, VP, I32, VP, VP, VP, VP]),
    "igm_transport_f32": (C.c_int, [I32, I32, VP, VP, VP, VP, VP, F64, F64, F64, I32, F64, F64, VP, VP, VP,
                                    VP, VP]),
    "igm_absmax2_f32": (C.c_int, [VP, VP, I64, VP, VP, VP]),
    "igm_smb_simple_f32": (C.c_int, [I64, VP, F64, F64, F64, F64, VP, VP]),
}

_lib = None


class IgmError(RuntimeError):
    pass


def load(path=None):
    """Load (once) and return the library with typed entry points."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise IgmError(f"{p} is missing: build it with `python -m igm_amd.build` "
                       "(the HIP library has no fallback)")
    lib = C.CDLL(p)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def declared_symbols():
    return list(_SIGS)


def check(rc):
    if rc != 0:
        msg = load().igm_last_error()
        raise IgmError(msg.decode() if msg else f"error code {rc}")


def ptr(t):
    """Device (or host) address of a tensor, or None."""
    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def call(name, *args):
    check(getattr(load(), name)(*args))
