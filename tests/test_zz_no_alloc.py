"""The entry point that allocates nothing (mms_potential_exact_noalloc):
device arrays in and out, chunk partials in the library's file-scope
scratch, results equal to calc_potential_exact bit for bit; two calls in a
row (the second on a smaller grid) and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F = np.float32


def _say(*a):
    print("[gpu-test]", *a, flush=True)


class _Hip:
    """Device buffers for the test through the HIP runtime (the entry point
    under test takes device pointers; the test allocates them).  The runtime
    functions are looked up through the library's own handle, so they come
    from the HIP runtime the library is linked against even when another
    copy of it is loaded in the process (PyTorch ships one)."""

    def __init__(self, mg):
        h = C.CDLL(mg.lib()._name)
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        self.h = h
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def put(self, arr):
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1) == 0
        return p

    def get(self, p, n):
        out = np.empty(n, F)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _entry(mg):
    return mg.lib().mms_potential_exact_noalloc


def test_noalloc_matches_calc_potential_exact(mg):
    fn = _entry(mg)
    hip = _Hip(mg)
    try:
        for dims, d in (((32, 32, 3), (F(5e-9),) * 3), ((13, 11, 3), (F(5e-9), F(4e-9), F(3e-9)))):
            n = int(np.prod(dims))
            _say("no-alloc direct sum", dims)
            q = (np.random.default_rng(n).normal(size=n) * 1e14).astype(F)
            ref = np.zeros(n, F)
            assert mg.calc_potential_exact(q, *dims, *d, ref, 1, 0) == 0, mg.last_error()
            dq = hip.put(q)
            dp = hip.alloc(4 * n)
            assert hip.h.hipMemset(dp, 0xFF, 4 * n) == 0
            assert fn(dq, dp, *dims, *d, None) == 0, mg.last_error()
            got = hip.get(dp, n)
            assert np.array_equal(got, ref), dims
    finally:
        hip.free()


def test_noalloc_rejects_grids_beyond_the_scratch(mg):
    fn = _entry(mg)
    hip = _Hip(mg)
    try:
        dq, dp = hip.alloc(16), hip.alloc(16)
        assert fn(dq, dp, 2049, 1024, 1, F(5e-9), F(5e-9), F(5e-9), None) != 0
        assert "scratch" in mg.last_error()
        assert fn(None, dp, 4, 4, 1, F(5e-9), F(5e-9), F(5e-9), None) != 0
    finally:
        hip.free()
