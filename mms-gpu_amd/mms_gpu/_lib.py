"""ctypes binding of libmmsgpu.so (the gfx950 HIP library).

The library is built in-tree (``mms-gpu_amd/libmmsgpu.so``); the environment
variable MMS_GPU_LIB may name another build of it.  Loading fails loudly when
it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("MMS_GPU_LIB") or os.path.join(PKG_ROOT, "libmmsgpu.so")

_lib = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
I, F, D, VP = C.c_int, C.c_float, C.c_double, C.c_void_p


class FilmSpec(C.Structure):
    """Mirror of ``mms_film_spec``."""
    _fields_ = [("xdim", I), ("ydim", I), ("hext", F * 3)]


class MmsParams(C.Structure):
    """Mirror of ``mms_params`` in include/mms_gpu.h."""
    _fields_ = [("xdim", I), ("ydim", I), ("zdim", I),
                ("dx", F), ("dy", F), ("dz", F),
                ("solver", I), ("P", I), ("m2l_precision", I), ("leaf_width", I),
                ("mu_0", F), ("Ms", F), ("Aexch", F), ("alfa", F), ("gamma", F),
                ("demag", I), ("exchange", I), ("external", I),
                ("hext", F * 3), ("m2l_cols", I), ("direct_split", I)]


# symbol -> (restype, argtypes); every symbol declared in include/mms_gpu.h
SIGNATURES = {
    "calc_potential_exact": (I, [f32p, I, I, I, F, F, F, f32p, I, I]),
    "calc_potential_exact_gpu": (I, [f32p, I, I, I, F, F, F, f32p]),
    "fmm_calc": (I, [f32p, f32p, I, I, I, F, F, F, I, I, I]),
    "Hfield": (I, [f32p, f32p, VP, VP, F, I, I, I, F, F, F, I, F, F, F, I, I, I, I, I, I]),
    "rk4_step": (I, [F, F, C.POINTER(F), u8p, f32p, f32p, f32p, VP, VP, VP, VP, I, I, I, F, F, F, I,
                     F, F, F, F, F, I, I, I, I, I, I, I]),
    "time_marching": (I, [u8p, f32p, F, F, I, I, I, F, F, F, I, F, F, F, F, F, I, I, I, I, I,
                          C.c_char_p, I]),
    "mms_potential_exact_noalloc": (I, [VP, VP, I, I, I, F, F, F, VP]),
    "mms_create": (I, [C.POINTER(VP), C.POINTER(MmsParams), VP]),
    "mms_destroy": (I, [VP]),
    "mms_set_m": (I, [VP, f32p]),
    "mms_get_m": (I, [VP, f32p]),
    "mms_set_time": (I, [VP, D]),
    "mms_steps": (I, [VP, I, F, I]),
    "mms_sync": (I, [VP]),
    "mms_hfield": (I, [VP, f32p]),
    "mms_energy_torque": (I, [VP, C.POINTER(D), C.POINTER(D)]),
    "mms_potential_dev": (I, [VP, VP, VP]),
    "mms_potential": (I, [VP, f32p, f32p]),
    "mms_potential_probes_f64": (I, [VP, f32p, i64p, I, f64p]),
    "mms_autotune": (I, [VP, I, C.c_char_p, I]),
    "mms_get_shape": (I, [VP, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
    "mms_profile_begin": (I, [VP]),
    "mms_profile_end": (I, [VP]),
    "mms_profile_count": (I, [VP]),
    "mms_profile_name": (C.c_char_p, [VP, I]),
    "mms_profile_ms": (D, [VP, I]),
    "mms_profile_launches": (C.c_long, [VP, I]),
    "mms_surrogate_weight_count": (C.c_long, []),
    "mms_surrogate_init": (I, [VP, VP, F, F]),
    "mms_surrogate_potential": (I, [VP, f32p, f32p]),
    "mms_surrogate_potential_dev": (I, [VP, VP, VP]),
    "mms_surrogate_mode": (I, [VP, I, F]),
    "mms_surrogate_stats": (I, [VP, C.POINTER(D), C.POINTER(C.c_long), C.POINTER(C.c_long)]),
    "mms_surrogate_state": (I, [I, I, I, F, C.c_ulonglong, C.c_long, f32p, VP, C.POINTER(I)]),
    "mms_surrogate_trainer_create": (I, [C.POINTER(VP), I, I, I, F, F, I, VP]),
    "mms_surrogate_trainer_destroy": (I, [VP]),
    "mms_surrogate_trainer_generate": (I, [VP, I, C.c_ulonglong, I]),
    "mms_surrogate_trainer_add": (I, [VP, I, f32p, f32p]),
    "mms_surrogate_trainer_pairs": (C.c_long, [VP]),
    "mms_surrogate_trainer_get_pair": (I, [VP, C.c_long, VP, VP]),
    "mms_surrogate_trainer_gradient": (I, [VP, i64p, VP, C.POINTER(D)]),
    "mms_surrogate_trainer_step": (I, [VP, i64p, F, C.POINTER(D)]),
    "mms_surrogate_trainer_eval": (I, [VP, i64p, f64p]),
    "mms_surrogate_trainer_predict": (I, [VP, i64p, f32p]),
    "mms_surrogate_trainer_weights": (I, [VP, f32p]),
    "mms_batch_create": (I, [C.POINTER(VP), C.POINTER(MmsParams), I, VP, VP]),
    "mms_batch_destroy": (I, [VP]),
    "mms_batch_set_m": (I, [VP, f32p]),
    "mms_batch_get_m": (I, [VP, f32p]),
    "mms_batch_steps": (I, [VP, I, F, I]),
    "mms_batch_sync": (I, [VP]),
    "mms_batch_potential": (I, [VP, f32p, f32p]),
    "mms_batch_info": (I, [VP, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(C.c_long), C.POINTER(I)]),
    "mms_batch_autotune": (I, [VP, I, C.c_char_p, I]),
    "mms_batch_profile_begin": (I, [VP]),
    "mms_batch_profile_end": (I, [VP, f64p, np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")]),
    "mms_last_error": (C.c_char_p, []),
    "mms_version": (C.c_char_p, []),
}


def lib():
    """Load libmmsgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(make -C mms-gpu_amd); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().mms_last_error().decode()


class MmsError(RuntimeError):
    pass


def check(status, what):
    if status != 0:
        raise MmsError(f"{what} failed: {last_error()}")
