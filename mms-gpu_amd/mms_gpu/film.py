"""Film: one magnetic film resident on the GPU (wraps an ``mms_ctx``)."""
import ctypes as C

import numpy as np

from ._lib import MmsParams, check, lib

SOLVER_DIRECT, SOLVER_FMM, SOLVER_DIRECT_REFGPU = 0, 1, 2
(M2L_FP32, M2L_BF16, M2L_BF16X3, M2L_FP16, M2L_BF16X3_FP16NEAR, M2L_FP16X2, M2L_FP16X2_BF16X3NEAR,
 M2L_FP16X2_FP16X2NEAR) = 0, 1, 2, 3, 4, 5, 6, 7
EXT_NONE, EXT_RAMP, EXT_CONST = 0, 1, 2

MU_0 = np.float32(4 * np.pi * 1e-7)


class Film:
    """Device-resident film state with fixed-step RK4.

    Parameters mirror ``mms_params`` (include/mms_gpu.h).  ``material`` is a
    uint8 array of zdim*ydim*xdim cells (None = all magnetic).
    """

    def __init__(self, xdim, ydim, zdim, dx, dy, dz, *, solver=SOLVER_DIRECT, P=6,
                 m2l_precision=M2L_FP32, leaf_width=4, mu_0=MU_0, Ms=8.6e5, Aexch=1.3e-11,
                 alfa=0.5, gamma=2.211e5, demag=1, exchange=1, external=EXT_NONE,
                 hext=(0.0, 0.0, 0.0), material=None, m2l_cols=4, direct_split=1):
        p = MmsParams()
        p.xdim, p.ydim, p.zdim = xdim, ydim, zdim
        p.dx, p.dy, p.dz = dx, dy, dz
        p.solver, p.P, p.m2l_precision, p.leaf_width = solver, P, m2l_precision, leaf_width
        p.mu_0, p.Ms, p.Aexch, p.alfa, p.gamma = mu_0, Ms, Aexch, alfa, gamma
        p.demag, p.exchange, p.external = demag, exchange, external
        for i in range(3):
            p.hext[i] = hext[i]
        p.m2l_cols, p.direct_split = m2l_cols, direct_split
        self.params = p
        self.dims = (xdim, ydim, zdim)
        self.n = xdim * ydim * zdim
        self._ctx = C.c_void_p()
        mat = None
        if material is not None:
            self._mat = np.ascontiguousarray(material, dtype=np.uint8).ravel()
            if self._mat.size != self.n:
                raise ValueError("material size does not match the grid")
            mat = self._mat.ctypes.data_as(C.c_void_p)
        check(lib().mms_create(C.byref(self._ctx), C.byref(p), mat), "mms_create")

    def close(self):
        if self._ctx:
            lib().mms_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---------------------------------------------------------- state --
    def set_m(self, M):
        M = np.ascontiguousarray(M, dtype=np.float32).ravel()
        check(lib().mms_set_m(self._ctx, M), "mms_set_m")

    def get_m(self):
        out = np.zeros(3 * self.n, np.float32)
        check(lib().mms_get_m(self._ctx, out), "mms_get_m")
        return out

    def set_time(self, t):
        check(lib().mms_set_time(self._ctx, float(t)), "mms_set_time")

    # --------------------------------------------- learned surrogate --
    def surrogate_init(self, weights=None, charge_scale=0.0, potential_scale=0.0):
        """Load the charge -> potential network (weights: FP32 vector of
        mms_surrogate_weight_count() values, None = seeded untrained)."""
        w = None
        if weights is not None:
            self._sur_w = np.ascontiguousarray(weights, dtype=np.float32).ravel()
            if self._sur_w.size != lib().mms_surrogate_weight_count():
                raise ValueError("surrogate weight vector has the wrong size")
            w = self._sur_w.ctypes.data_as(C.c_void_p)
        check(lib().mms_surrogate_init(self._ctx, w, float(charge_scale), float(potential_scale)),
              "mms_surrogate_init")

    def surrogate_potential(self, charge):
        out = np.zeros(self.n, np.float32)
        check(lib().mms_surrogate_potential(self._ctx, np.ascontiguousarray(charge, np.float32).ravel(), out),
              "mms_surrogate_potential")
        return out

    def surrogate_mode(self, mode, tol=1e-3):
        check(lib().mms_surrogate_mode(self._ctx, int(mode), float(tol)), "mms_surrogate_mode")

    def surrogate_stats(self):
        r, j, a = C.c_double(), C.c_long(), C.c_long()
        check(lib().mms_surrogate_stats(self._ctx, C.byref(r), C.byref(j), C.byref(a)), "mms_surrogate_stats")
        return {"last_residual": r.value, "judged": j.value, "accepted": a.value}

    # ------------------------------------------------------- stepping --
    def steps(self, nsteps, dt, use_graph=True):
        check(lib().mms_steps(self._ctx, int(nsteps), float(dt), int(bool(use_graph))), "mms_steps")

    def sync(self):
        check(lib().mms_sync(self._ctx), "mms_sync")

    # --------------------------------------------------------- fields --
    def hfield(self):
        out = np.zeros(3 * self.n, np.float32)
        check(lib().mms_hfield(self._ctx, out), "mms_hfield")
        return out

    def energy_torque(self):
        E, T = C.c_double(), C.c_double()
        check(lib().mms_energy_torque(self._ctx, C.byref(E), C.byref(T)), "mms_energy_torque")
        return E.value, T.value

    def potential(self, charge):
        charge = np.ascontiguousarray(charge, dtype=np.float32).ravel()
        out = np.zeros(self.n, np.float32)
        check(lib().mms_potential(self._ctx, charge, out), "mms_potential")
        return out

    def potential_probes_f64(self, charge, probes):
        charge = np.ascontiguousarray(charge, dtype=np.float32).ravel()
        probes = np.ascontiguousarray(probes, dtype=np.int64).ravel()
        out = np.zeros(len(probes), np.float64)
        check(lib().mms_potential_probes_f64(self._ctx, charge, probes, len(probes), out),
              "mms_potential_probes_f64")
        return out

    # ------------------------------------------------------- autotuner --
    def autotune(self, reps=3):
        """Time the solver's launch shapes on the current state's charge and
        keep the fastest; returns the per-candidate report."""
        buf = C.create_string_buffer(4096)
        check(lib().mms_autotune(self._ctx, int(reps), buf, 4096), "mms_autotune")
        return buf.value.decode()

    def shape(self):
        w, c, s = C.c_int(), C.c_int(), C.c_int()
        check(lib().mms_get_shape(self._ctx, C.byref(w), C.byref(c), C.byref(s)), "mms_get_shape")
        return {"leaf_width": w.value, "m2l_cols": c.value, "direct_split": s.value}

    # -------------------------------------------------------- profiling --
    def profile_begin(self):
        check(lib().mms_profile_begin(self._ctx), "mms_profile_begin")

    def profile_end(self):
        """Stop profiling; returns {kernel class: (mean ms per launch, launches)}."""
        check(lib().mms_profile_end(self._ctx), "mms_profile_end")
        out = {}
        for i in range(lib().mms_profile_count(self._ctx)):
            cnt = lib().mms_profile_launches(self._ctx, i)
            if cnt:
                out[lib().mms_profile_name(self._ctx, i).decode()] = (
                    lib().mms_profile_ms(self._ctx, i), cnt)
        return out
