"""ctypes binding of liblenia_hip.so (the C-ABI declared in include/lenia_hip.h).

The library is loaded from this package directory only; a missing library is an
error, never a silent fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblenia_hip.so")

MAX_K, MAX_C, MAX_BETA = 16, 4, 12

OK, EINVAL, EUNSUPPORTED, EHIP, ESTATE, ENOMEM = 0, -1, -2, -3, -4, -5
F_A, F_GA, F_S, F_V, F_H, F_B, F_MU, F_USUM = range(8)
OPT_KEEP_VELOCITY, OPT_KEEP_USUM, OPT_USE_GRAPH, OPT_PROFILE, OPT_AUTOTUNE, OPT_SCHEDULE, OPT_KEEP_MU = range(7)

MODE = {"srt": 0, "rt": 1, "adv": 2}
BORDER = {"torus": 0, "wall": 1, "none": 2}
GRADIENT = {"central": 0, "sobel": 1}
GROWTH = {"none": 0, "rt": 1, "transposition": 2}
Q_ALL = 0x7FF

# Every symbol include/lenia_hip.h declares.
EXPORTS = ("lh_create", "lh_destroy", "lh_set_option", "lh_add_world", "lh_commit", "lh_set_field",
           "lh_get_field", "lh_field_device_ptr", "lh_state_updated", "lh_step", "lh_prepare_steps", "lh_synchronize",
           "lh_stream", "lh_world_mass", "lh_mc_case", "lh_mc_case_noalloc", "lh_mc_table", "lh_mesh", "lh_mesh_snapshot", "lh_world_count", "lh_profile_count",
           "lh_profile_entry", "lh_profile_reset", "lh_calibration_copy", "lh_shape_supported", "lh_last_error", "lh_version",
           "lh_set_world_conv", "lh_world_moments",
           "lh_checkpoint_size", "lh_checkpoint_save", "lh_checkpoint_load", "lh_world_steps",
           "lh_sample_random", "lh_seed_random", "lh_snapshot", "lh_snapshot_ready", "lh_snapshot_read",
           "lm_param_count", "lm_create", "lm_destroy", "lm_set_params", "lm_get_params", "lm_set_data",
           "lm_train", "lm_gradients", "lm_forward", "lm_score_sampled", "lm_sample_features")

# growth convolution methods (lh_set_world_conv)
CONV = {"fft": 0, "bf16x3": 1, "fp16x3": 2, "bf16": 3, "fp16": 4}


class Params(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("border", C.c_int), ("sobel_border", C.c_int),
        ("gradient", C.c_int), ("growth_type", C.c_int),
        ("mu_vel", C.c_int), ("adv_vel", C.c_int), ("blur", C.c_int),
        ("dt", C.c_float), ("sigma", C.c_float), ("dd", C.c_float), ("s_norm", C.c_float),
        ("theta_a", C.c_float), ("gta", C.c_float), ("sta", C.c_float), ("v_damp", C.c_float),
        ("asymptotic", C.c_int * MAX_C), ("soft_clip", C.c_int * MAX_C),
        ("osc_la", C.c_int), ("osc_flow", C.c_int), ("osc_div_flow", C.c_int), ("osc_vel", C.c_int),
        ("osc_la_amt", C.c_float), ("osc_flow_amt", C.c_float), ("osc_div_flow_amt", C.c_float),
        ("quirks", C.c_uint32),
    ]


class Kernel(C.Structure):
    _fields_ = [
        ("c0", C.c_int), ("c1", C.c_int), ("radius", C.c_int), ("beta_size", C.c_int),
        ("b", C.c_float * MAX_BETA), ("r_a", C.c_float), ("m", C.c_float), ("s", C.c_float),
        ("h", C.c_float),
    ]


class Shape(C.Structure):
    _fields_ = [("W", C.c_int), ("H", C.c_int), ("D", C.c_int), ("C", C.c_int), ("K", C.c_int)]


CHECKPOINT_VERSION = 1


class CheckpointHeader(C.Structure):
    """lh_checkpoint_header (64 bytes)."""
    _fields_ = [
        ("magic", C.c_char * 4), ("version", C.c_uint32), ("header_bytes", C.c_uint32),
        ("params_bytes", C.c_uint32), ("kernel_bytes", C.c_uint32), ("conv", C.c_int32),
        ("shape", Shape), ("reserved", C.c_uint32), ("steps", C.c_uint64), ("payload_bytes", C.c_uint64),
    ]


class LeniaHipError(RuntimeError):
    pass


_lib = None


def load():
    """Loads liblenia_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LeniaHipError(f"{LIB_PATH} is missing: run `python -m lenia_amd.build` "
                            "(or __graft_entry__.build()) first")
    L = C.CDLL(LIB_PATH)
    vp, i, f, st = C.c_void_p, C.c_int, C.c_float, C.c_size_t
    sig = {
        "lh_create": (i, [i, C.POINTER(vp)]),
        "lh_destroy": (i, [vp]),
        "lh_set_option": (i, [vp, i, i]),
        "lh_add_world": (i, [vp, C.POINTER(Shape), C.POINTER(Params), C.POINTER(Kernel), C.POINTER(i)]),
        "lh_commit": (i, [vp]),
        "lh_set_world_conv": (i, [vp, i, i]),
        "lh_set_field": (i, [vp, i, i, C.POINTER(f), st]),
        "lh_get_field": (i, [vp, i, i, C.POINTER(f), st]),
        "lh_field_device_ptr": (i, [vp, i, i, C.POINTER(C.POINTER(f)), C.POINTER(st)]),
        "lh_state_updated": (i, [vp]),
        "lh_step": (i, [vp, i]),
        "lh_prepare_steps": (i, [vp, i]),
        "lh_synchronize": (i, [vp]),
        "lh_stream": (vp, [vp]),
        "lh_world_mass": (i, [vp, C.POINTER(C.c_double), i]),
        "lh_world_moments": (i, [vp, C.POINTER(C.c_double), i]),
        "lh_checkpoint_size": (i, [vp, i, C.POINTER(st)]),
        "lh_checkpoint_save": (i, [vp, i, vp, st]),
        "lh_checkpoint_load": (i, [vp, i, vp, st]),
        "lh_world_steps": (i, [vp, i, C.POINTER(C.c_uint64)]),
        "lh_sample_random": (i, [i, C.c_uint64, C.c_int64, i, i, C.POINTER(Kernel), C.POINTER(i)]),
        "lh_seed_random": (i, [vp, i, i, C.c_uint64, C.c_int64]),
        "lh_snapshot": (i, [vp, i, i, i, i]),
        "lh_snapshot_ready": (i, [vp, i, C.POINTER(i)]),
        "lh_snapshot_read": (i, [vp, i, i, i, C.POINTER(f), st]),
        "lh_mc_case": (i, [vp, i, i, i, f, i, C.POINTER(C.c_uint8)]),
        "lh_mc_case_noalloc": (i, [vp, i, i, i, f, i, C.POINTER(C.c_uint8)]),
        "lh_mc_table": (i, [C.POINTER(i)]),
        "lh_mesh": (i, [vp, i, i, f, f, i, C.POINTER(i), C.POINTER(C.c_uint32), vp, C.c_int64, C.POINTER(C.c_uint32),
                        C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "lh_mesh_snapshot": (i, [vp, i, i, i, f, f, i, C.POINTER(i), C.POINTER(C.c_uint32), vp, C.c_int64,
                                 C.POINTER(C.c_uint32), C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "lh_world_count": (i, [vp]),
        "lh_profile_count": (i, [vp]),
        "lh_profile_entry": (i, [vp, i, C.POINTER(C.c_char_p), C.POINTER(C.c_long), C.POINTER(C.c_double)]),
        "lh_profile_reset": (i, [vp]),
        "lh_calibration_copy": (i, [vp, st]),
        "lh_shape_supported": (i, [i, i, i]),
        "lh_last_error": (C.c_char_p, []),
        "lm_param_count": (st, []),
        "lm_create": (i, [i, i, C.POINTER(vp)]),
        "lm_destroy": (i, [vp]),
        "lm_set_params": (i, [vp, C.POINTER(f), st]),
        "lm_get_params": (i, [vp, C.POINTER(f), st]),
        "lm_set_data": (i, [vp, C.POINTER(f), C.POINTER(f), i]),
        "lm_train": (i, [vp, C.POINTER(i), i, i, f, C.POINTER(f)]),
        "lm_gradients": (i, [vp, C.POINTER(f), C.POINTER(f), i, C.POINTER(f), C.POINTER(f)]),
        "lm_forward": (i, [vp, C.POINTER(f), i, C.POINTER(f)]),
        "lm_score_sampled": (i, [vp, C.c_uint64, C.c_longlong, C.POINTER(f)]),
        "lm_sample_features": (i, [vp, C.c_uint64, C.c_longlong, i, C.POINTER(f)]),
        "lh_version": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    _lib = L
    return L


def check(rc: int, what: str = ""):
    if rc != OK:
        msg = load().lh_last_error().decode(errors="replace")
        raise LeniaHipError(f"{what} failed ({rc}): {msg}")


def make_params(cfg: dict) -> Params:
    p = Params()
    p.mode = MODE[cfg.get("mode", "rt")]
    p.border = BORDER[cfg.get("border", "torus")]
    p.sobel_border = BORDER[cfg.get("sobel_border", "torus")]
    p.gradient = GRADIENT[cfg.get("gradient", "sobel")]
    p.growth_type = GROWTH[cfg.get("growth", "rt")]
    p.mu_vel = int(cfg.get("mu_vel", False))
    p.adv_vel = int(cfg.get("adv_vel", True))
    p.blur = int(cfg.get("blur", False))
    for k, d in (("dt", 0.2), ("sigma", 0.65), ("dd", 1.0), ("s_norm", 0.0), ("theta_a", 2.0),
                 ("gta", 0.1), ("sta", 0.1), ("v_damp", 0.9), ("osc_la_amt", -0.001),
                 ("osc_flow_amt", -0.001), ("osc_div_flow_amt", -0.001)):
        setattr(p, k, float(cfg.get(k, d)))
    for c in range(MAX_C):
        p.asymptotic[c] = int(cfg.get("asymptotic", [0] * MAX_C)[c])
        p.soft_clip[c] = int(cfg.get("soft_clip", [0] * MAX_C)[c])
    for k in ("osc_la", "osc_flow", "osc_div_flow", "osc_vel"):
        setattr(p, k, int(cfg.get(k, False)))
    p.quirks = int(cfg.get("quirks", Q_ALL))
    return p


def make_kernels(kernels: list):
    arr = (Kernel * max(1, len(kernels)))()
    for j, k in enumerate(kernels):
        arr[j].c0, arr[j].c1 = int(k["c0"]), int(k["c1"])
        arr[j].radius, arr[j].beta_size = int(k["radius"]), int(k["beta_size"])
        for q, b in enumerate(k["b"]):
            arr[j].b[q] = float(b)
        arr[j].r_a, arr[j].m, arr[j].s, arr[j].h = (float(k[n]) for n in ("r_a", "m", "s", "h"))
    return arr
