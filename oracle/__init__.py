"""CPU oracle for the field-evaluation + LLG RK4 path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s ``cpu_baseline`` leg, and nothing else.  The product
library never routes through it.

* ``lib()``      -- ctypes handle on ``oracle/liboracle.so`` (mms_oracle.c:
                    FP32 restatements of the reference CPU path + FP64 direct sum)
* ``ref_lib()``  -- ctypes handle on ``oracle/_ref/libmmsref.so`` (the reference's
                    own vector_functions / numerics / Box / LLG_Mprime, compiled
                    from its sources), or None when it was not built
* ``fmm_oracle`` -- numpy FP64 regular-grid FMM and FFT-convolution direct sum
"""
import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref = None

f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
I, F, D = C.c_int, C.c_float, C.c_double


def _sig(lib, name, res, args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        _sig(L, "or_associated_legendre", F, [I, I, F])
        _sig(L, "or_spherical_harmonic", None, [I, I, F, F, C.POINTER(F), C.POINTER(F)])
        _sig(L, "or_divergence_3d", None, [f32p, I, I, I, F, F, F, f32p])
        _sig(L, "or_gradient_3d", None, [f32p, I, I, I, F, F, F, F, f32p])
        _sig(L, "or_add_exchange_field", None, [f32p, F, F, F, I, I, I, F, F, F, f32p])
        _sig(L, "or_calc_potential_exact", I, [f32p, I, I, I, F, F, F, f32p])
        _sig(L, "or_calc_potential_exact_refgpu", I, [f32p, I, I, I, F, F, F, f32p])
        _sig(L, "or_potential_exact_f64", None, [f32p, I, I, I, D, D, D, f64p])
        _sig(L, "or_potential_probes_f64", None, [f32p, I, I, I, D, D, D, i64p, I, f64p])
        _sig(L, "or_tree_lists", None, [I, i32p, i32p])
        _sig(L, "or_fmm_calc", I, [f32p, f32p, I, I, I, F, F, F, I])
        _sig(L, "or_fmm_bfs_sample", I, [f32p, f32p, I, I, I, F, F, F, I, I, I, I, I, I])
        _sig(L, "or_llg_mprime", None, [f32p, f32p, F, F, F, f32p])
        _sig(L, "or_hfield", I, [f32p, f32p, f32p, f32p, F, I, I, I, F, F, F, I, F, F, F,
                                 I, I, I, I, C.c_void_p])
        _sig(L, "or_rk4_step", I, [F, F, C.POINTER(F), u8p, f32p, f32p, f32p, f32p, f32p,
                                   f32p, f32p, I, I, I, F, F, F, I, F, F, F, F, F,
                                   I, I, I, I, I, C.c_void_p, I])
        _sig(L, "or_energy_torque", None, [u8p, f32p, f32p, I, F, F, F, F, F,
                                           C.POINTER(F), C.POINTER(F)])
        _sig(L, "or_num_threads", I, [])
        _lib = L
    return _lib


def ref_lib():
    """The reference's own CPU functions (oracle/_ref), or None if absent."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libmmsref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        _sig(R, "ref_divergence_3d", None, [f32p, I, I, I, F, F, F, f32p])
        _sig(R, "ref_gradient_3d", None, [f32p, I, I, I, F, F, F, F, f32p])
        _sig(R, "ref_add_exchange_field", None, [f32p, F, F, F, I, I, I, F, F, F, f32p])
        _sig(R, "ref_associated_legendre", F, [I, I, F])
        _sig(R, "ref_spherical_harmonic", None, [I, I, F, F, C.POINTER(F), C.POINTER(F)])
        _sig(R, "ref_factorial", I, [I])
        _sig(R, "ref_llg_mprime", None, [f32p, f32p, F, F, F, f32p])
        _sig(R, "ref_tree_lists", I, [I, I, i32p, i32p, f32p])
        _ref = R
    return _ref


# ---------------------------------------------------------------- wrappers --
def _c(a, dt=np.float32):
    return np.ascontiguousarray(a, dtype=dt)


def divergence_3d(M, dims, d):
    xd, yd, zd = dims
    S = np.zeros(xd * yd * zd, np.float32)
    lib().or_divergence_3d(_c(M).ravel(), xd, yd, zd, *d, S)
    return S


def gradient_3d(S, dims, d, cm):
    xd, yd, zd = dims
    V = np.zeros(3 * xd * yd * zd, np.float32)
    lib().or_gradient_3d(_c(S).ravel(), xd, yd, zd, *d, cm, V)
    return V


def add_exchange_field(M, H, Ms, A, mu, dims, d):
    xd, yd, zd = dims
    H = _c(H).ravel().copy()
    lib().or_add_exchange_field(_c(M).ravel(), Ms, A, mu, xd, yd, zd, *d, H)
    return H


def calc_potential_exact(charge, dims, d):
    xd, yd, zd = dims
    pot = np.zeros(xd * yd * zd, np.float32)
    lib().or_calc_potential_exact(_c(charge).ravel(), xd, yd, zd, *d, pot)
    return pot


def calc_potential_exact_refgpu(charge, dims, d):
    """The reference CUDA kernel's order (potential_calc.cu:18-116), FP32."""
    xd, yd, zd = dims
    pot = np.zeros(xd * yd * zd, np.float32)
    lib().or_calc_potential_exact_refgpu(_c(charge).ravel(), xd, yd, zd, *d, pot)
    return pot


def potential_exact_f64(charge, dims, d):
    xd, yd, zd = dims
    pot = np.zeros(xd * yd * zd, np.float64)
    lib().or_potential_exact_f64(_c(charge).ravel(), xd, yd, zd, *[float(v) for v in d], pot)
    return pot


def potential_probes_f64(charge, dims, d, probes):
    xd, yd, zd = dims
    probes = _c(probes, np.int64).ravel()
    out = np.zeros(len(probes), np.float64)
    lib().or_potential_probes_f64(_c(charge).ravel(), xd, yd, zd, *[float(v) for v in d],
                                  probes, len(probes), out)
    return out


def fmm_calc(charge, dims, d, P):
    xd, yd, zd = dims
    pot = np.zeros(xd * yd * zd, np.float32)
    st = lib().or_fmm_calc(_c(charge).ravel(), pot, xd, yd, zd, *d, P)
    if st:
        raise ValueError("reference tree-code needs xdim == ydim == 2^k and P <= 4")
    return pot


def hfield(M, dims, d, P, mu_0, Ms, A, demag=1, exchange=1, external=0, use_fmm=0, t=0.0,
           hext=None):
    xd, yd, zd = dims
    n = xd * yd * zd
    H = np.zeros(3 * n, np.float32)
    ch = np.zeros(n, np.float32)
    pot = np.zeros(n, np.float32)
    hp = None
    if hext is not None:
        hbuf = (C.c_float * 3)(*[float(v) for v in hext])
        hp = C.cast(hbuf, C.c_void_p)
    st = lib().or_hfield(_c(M).ravel(), H, ch, pot, t, xd, yd, zd, *d, P, mu_0, Ms, A,
                         demag, exchange, external, use_fmm, hp)
    if st:
        raise RuntimeError("or_hfield failed")
    return H, ch, pot


def rk4_step(M, H, material, dims, d, dt, P, mu_0, Ms, A, alfa, gamma, demag=1, exchange=1,
             external=0, use_fmm=0, normalize=1, t=0.0, hext=None, combine64=0):
    """One reference rk4_step.  H is the field at M (k1); returns (M2, H_last).
    M2 starts as a copy of M (non-material cells keep it)."""
    M2, Hb, _, _ = rk4_step_full(M, H, material, dims, d, dt, P, mu_0, Ms, A, alfa, gamma, demag,
                                 exchange, external, use_fmm, normalize, t, hext, combine64)
    return M2, Hb


def rk4_step_full(M, H, material, dims, d, dt, P, mu_0, Ms, A, alfa, gamma, demag=1, exchange=1,
                  external=0, use_fmm=0, normalize=1, t=0.0, hext=None, combine64=0, M2_in=None,
                  this_slope_in=None):
    """One reference rk4_step with every output: returns (M2, H_last, slope,
    this_slope).  M2_in / this_slope_in are the caller's arrays on entry (the
    reference leaves them untouched on non-material cells); default: a copy
    of M and zeros."""
    xd, yd, zd = dims
    n = xd * yd * zd
    M = _c(M).ravel()
    M2 = M.copy() if M2_in is None else _c(M2_in).ravel().copy()
    Hb = _c(H).ravel().copy()
    slope = np.zeros(3 * n, np.float32)
    ts = np.zeros(3 * n, np.float32) if this_slope_in is None else _c(this_slope_in).ravel().copy()
    ch = np.zeros(n, np.float32)
    pot = np.zeros(n, np.float32)
    t2 = C.c_float(0)
    hp = None
    if hext is not None:
        hbuf = (C.c_float * 3)(*[float(v) for v in hext])
        hp = C.cast(hbuf, C.c_void_p)
    lib().or_rk4_step(t, dt, C.byref(t2), _c(material, np.uint8).ravel(), M, M2, Hb, slope, ts,
                      ch, pot, xd, yd, zd, *d, P, mu_0, Ms, A, alfa, gamma, demag, exchange,
                      external, use_fmm, normalize, hp, combine64)
    return M2, Hb, slope, ts


def energy_torque(material, M, H, mu_0, Ms, d):
    E = C.c_float(0)
    T = C.c_float(0)
    material = _c(material, np.uint8).ravel()
    lib().or_energy_torque(material, _c(M).ravel(), _c(H).ravel(), len(material), mu_0, Ms,
                           *d, C.byref(E), C.byref(T))
    return E.value, T.value


def tree_lists(level):
    nb = 4 ** level
    neighbors = np.zeros(nb * 8, np.int32)
    inter = np.zeros(nb * 27, np.int32)
    lib().or_tree_lists(level, neighbors, inter)
    return neighbors.reshape(nb, 8), inter.reshape(nb, 27)
