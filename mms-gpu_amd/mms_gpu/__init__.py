"""mms_gpu -- MI355X (gfx950) field evaluation and LLG RK4 time stepping.

Python mirror of the reference's entry points (same names, argument order
and meaning, 0/1 status returns, in-place output arrays):

    calc_potential_exact(charge, xdim, ydim, zdim, dx, dy, dz, potential, use_gpu, verbosity)
    calc_potential_exact_gpu(charge, xdim, ydim, zdim, dx, dy, dz, potential)
    fmm_calc(charge, potential, xdim, ydim, zdim, dx, dy, dz, P, use_gpu, verbosity)
    Hfield(M, H, charge, potential, t, xdim, ydim, zdim, dx, dy, dz, P, mu_0, Ms, Aexch,
           demag, exchange, external, use_fmm, use_gpu, verbosity)
    rk4_step(t, dt, material, M, M2, H, slope, this_slope, charge, potential, xdim, ydim, zdim,
             dx, dy, dz, P, mu_0, Ms, Aexch, alfa, gamma, demag, exchange, external, use_fmm,
             normalize, use_gpu, verbosity) -> (status, t2)
    time_marching(material, M, finaltime, timestep, xdim, ydim, zdim, dx, dy, dz, P, mu_0, Ms,
                  Aexch, alfa, gamma, demag, exchange, external, use_fmm, use_gpu, sim_name,
                  verbosity)

Arrays are numpy: scalars float32 [zdim*ydim*xdim], vector fields float32
[zdim*ydim*xdim*3] (x, y, z per cell, the reference's Vector3 layout),
material uint8.  All work runs on the GPU through libmmsgpu.so.

``Film`` is the resident interface (device state, fixed-step RK4, solver and
M2L precision selection, timing).
"""
import ctypes as C

import numpy as np

from ._lib import MmsError, MmsParams, check, last_error, lib  # noqa: F401
from .film import Film, SOLVER_DIRECT, SOLVER_FMM, M2L_FP32, M2L_BF16, M2L_BF16X3, M2L_FP16, M2L_BF16X3_FP16NEAR, M2L_FP16X2  # noqa: F401
from .film import M2L_FP16X2_BF16X3NEAR, M2L_FP16X2_FP16X2NEAR, SOLVER_DIRECT_REFGPU  # noqa: F401
from .film import EXT_NONE, EXT_RAMP, EXT_CONST  # noqa: F401
from .batch import FilmBatch  # noqa: F401
from .surrogate import SurrogateTrainer, surrogate_state  # noqa: F401


def _f32(a, name):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous):
        raise TypeError(f"{name} must be a C-contiguous float32 numpy array")
    return a


def _opt(a, name):
    if a is None:
        return None
    return _f32(a, name).ctypes.data_as(C.c_void_p)


def calc_potential_exact(charge, xdim, ydim, zdim, dx, dy, dz, potential, use_gpu=1, verbosity=0):
    """fmm_calc.cu:412-461 -- O(N^2) potential, fully overwrites `potential`."""
    return lib().calc_potential_exact(_f32(charge, "charge"), xdim, ydim, zdim, dx, dy, dz,
                                      _f32(potential, "potential"), use_gpu, verbosity)


def calc_potential_exact_gpu(charge, xdim, ydim, zdim, dx, dy, dz, potential):
    """potential_calc.cu:120-188 -- same computation as calc_potential_exact."""
    return lib().calc_potential_exact_gpu(_f32(charge, "charge"), xdim, ydim, zdim, dx, dy, dz,
                                          _f32(potential, "potential"))


def fmm_calc(charge, potential, xdim, ydim, zdim, dx, dy, dz, P, use_gpu=1, verbosity=0):
    """fmm_calc.cu:295-408 -- order-P fast multipole potential."""
    return lib().fmm_calc(_f32(charge, "charge"), _f32(potential, "potential"), xdim, ydim, zdim,
                          dx, dy, dz, P, use_gpu, verbosity)


def Hfield(M, H, charge, potential, t, xdim, ydim, zdim, dx, dy, dz, P, mu_0, Ms, Aexch,
           demag, exchange, external, use_fmm, use_gpu=1, verbosity=0):
    """ode_functions.cpp:21-70 -- effective field of M into H."""
    return lib().Hfield(_f32(M, "M"), _f32(H, "H"), _opt(charge, "charge"),
                        _opt(potential, "potential"), t, xdim, ydim, zdim, dx, dy, dz, P, mu_0,
                        Ms, Aexch, demag, exchange, external, use_fmm, use_gpu, verbosity)


def rk4_step(t, dt, material, M, M2, H, slope, this_slope, charge, potential, xdim, ydim, zdim,
             dx, dy, dz, P, mu_0, Ms, Aexch, alfa, gamma, demag, exchange, external, use_fmm,
             normalize=1, use_gpu=1, verbosity=0):
    """ode_functions.cpp:76-148 -- one RK4 step; returns (status, t2)."""
    t2 = C.c_float(0.0)
    mat = np.ascontiguousarray(material, dtype=np.uint8)
    st = lib().rk4_step(t, dt, C.byref(t2), mat, _f32(M, "M"), _f32(M2, "M2"), _f32(H, "H"),
                        _opt(slope, "slope"), _opt(this_slope, "this_slope"),
                        _opt(charge, "charge"), _opt(potential, "potential"), xdim, ydim, zdim,
                        dx, dy, dz, P, mu_0, Ms, Aexch, alfa, gamma, demag, exchange, external,
                        use_fmm, normalize, use_gpu, verbosity)
    return st, t2.value


def time_marching(material, M, finaltime, timestep, xdim, ydim, zdim, dx, dy, dz, P, mu_0, Ms,
                  Aexch, alfa, gamma, demag, exchange, external, use_fmm, use_gpu=1,
                  sim_name="sim_untitled", verbosity=0):
    """ode_functions.cpp:158-365 -- adaptive time loop; M is overwritten."""
    mat = np.ascontiguousarray(material, dtype=np.uint8)
    return lib().time_marching(mat, _f32(M, "M"), finaltime, timestep, xdim, ydim, zdim, dx, dy,
                               dz, P, mu_0, Ms, Aexch, alfa, gamma, demag, exchange, external,
                               use_fmm, use_gpu, sim_name.encode(), verbosity)


def version():
    return lib().mms_version().decode()
