"""FilmBatch: many independent films advanced together on one GPU."""
import ctypes as C

import numpy as np

from ._lib import FilmSpec, MmsParams, check, lib
from .film import MU_0, SOLVER_FMM, M2L_FP32


class FilmBatch:
    """Wraps ``mms_batch``.  ``films`` is a list of (xdim, ydim, hext) or of
    objects with xdim/ydim/hext/material attributes; materials are packed in
    film order (None = all magnetic)."""

    def __init__(self, films, zdim, dx, dy, dz, *, solver=SOLVER_FMM, P=4, m2l_precision=M2L_FP32,
                 leaf_width=4, m2l_cols=4, mu_0=MU_0, Ms=8.6e5, Aexch=1.3e-11, alfa=0.5, gamma=2.211e5,
                 demag=1, exchange=1, material=None):
        p = MmsParams()
        p.zdim, p.dx, p.dy, p.dz = zdim, dx, dy, dz
        p.solver, p.P, p.m2l_precision, p.leaf_width, p.m2l_cols = solver, P, m2l_precision, leaf_width, m2l_cols
        p.mu_0, p.Ms, p.Aexch, p.alfa, p.gamma = mu_0, Ms, Aexch, alfa, gamma
        p.demag, p.exchange = demag, exchange
        specs = (FilmSpec * len(films))()
        self.sizes = []
        for i, f in enumerate(films):
            xd, yd, h = (f.xdim, f.ydim, f.hext) if hasattr(f, "xdim") else f
            specs[i].xdim, specs[i].ydim = xd, yd
            for k in range(3):
                specs[i].hext[k] = h[k]
            self.sizes.append(xd * yd * zdim)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.ntot = int(self.offsets[-1])
        mat = None
        if material is not None:
            self._mat = np.ascontiguousarray(material, dtype=np.uint8).ravel()
            if self._mat.size != self.ntot:
                raise ValueError("material size does not match the films")
            mat = self._mat.ctypes.data_as(C.c_void_p)
        self._specs = specs
        self._b = C.c_void_p()
        check(lib().mms_batch_create(C.byref(self._b), C.byref(p), len(films), C.cast(specs, C.c_void_p), mat),
              "mms_batch_create")

    def close(self):
        if self._b:
            lib().mms_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_m(self, M):
        check(lib().mms_batch_set_m(self._b, np.ascontiguousarray(M, np.float32).ravel()), "mms_batch_set_m")

    def get_m(self):
        out = np.zeros(3 * self.ntot, np.float32)
        check(lib().mms_batch_get_m(self._b, out), "mms_batch_get_m")
        return out

    def film_m(self, M, f):
        return M[3 * self.offsets[f]:3 * self.offsets[f + 1]]

    def steps(self, nsteps, dt, use_graph=True):
        check(lib().mms_batch_steps(self._b, int(nsteps), float(dt), int(bool(use_graph))), "mms_batch_steps")

    def sync(self):
        check(lib().mms_batch_sync(self._b), "mms_batch_sync")

    def potential(self, charge):
        """Potential of every film (packed host arrays, film order) through
        the batch's own solvers."""
        out = np.zeros(self.ntot, np.float32)
        check(lib().mms_batch_potential(self._b, np.ascontiguousarray(charge, np.float32).ravel(), out),
              "mms_batch_potential")
        return out

    def info(self):
        a, b, c, e = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        d = C.c_long()
        check(lib().mms_batch_info(self._b, C.byref(a), C.byref(b), C.byref(c), C.byref(d), C.byref(e)),
              "mms_batch_info")
        return {"direct_films": a.value, "fmm_films": b.value, "fmm_size_classes": c.value, "cells": d.value,
                "launches_per_step": e.value}

    def autotune(self, reps=3):
        buf = C.create_string_buffer(8192)
        check(lib().mms_batch_autotune(self._b, int(reps), buf, 8192), "mms_batch_autotune")
        return buf.value.decode()

    def profile_begin(self):
        check(lib().mms_batch_profile_begin(self._b), "mms_batch_profile_begin")

    def profile_end(self):
        n = lib().mms_profile_count(None)
        ms = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int64)
        check(lib().mms_batch_profile_end(self._b, ms, cnt), "mms_batch_profile_end")
        return {lib().mms_profile_name(None, i).decode(): (ms[i], int(cnt[i])) for i in range(n) if cnt[i]}
