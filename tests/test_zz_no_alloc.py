"""rsg_cov_Calculate_noalloc (the allocation-free scoring entry point) through the C-ABI against
rsg_cov_Calculate: identical COV and extremes on two calls in a row, the second on a smaller
alignment, and a clear error beyond the fixed scratch size.  Device buffers for the input codes
and the output matrix come from the HIP runtime (hipMalloc through ctypes)."""
import ctypes as C

import numpy as np
import pytest

import rscape_amd as R
from rscape_amd import synth

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, nbytes):
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), nbytes) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.n
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0

    def get(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def _declare():
    L = R.lib()
    f = L.rsg_cov_Calculate_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                  C.c_int, C.c_double, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_char_p]
    return f


def noalloc(ctx, d_ax, N, L, d_cov, covtype, prec):
    f = _declare()
    mn, mx = C.c_double(), C.c_double()
    err = C.create_string_buffer(128)
    rc = f(ctx.h, d_ax.p, N, L, R.COV[covtype], R.CSELECT, prec, None, 8, 50, 1e-6, d_cov.p, C.byref(mn),
           C.byref(mx), err)
    return rc, mn.value, mx.value, err.value.decode()


def test_noalloc_matches_cov_calculate(ctx):
    cases = [(1500, 700, "GTp", R.STAT_FP64, 1), (200, 120, "MIa", R.STAT_FP32, 2)]   # the second is smaller
    d_ax = DevBuf(1500 * 700)
    d_cov = DevBuf(8 * 700 * 700)
    try:
        for N, L, covtype, prec, seed in cases:
            print("no-alloc case N=%d L=%d %s prec %d" % (N, L, covtype, prec), flush=True)
            ax = synth.config(0, seed=seed, N=N, L=L)["ax"]
            m = ctx.msa(ax)
            mn0, mx0 = m.calculate(covtype, prec=prec)
            ref = m.get_cov()[0]
            m.close()
            d_ax.put(ax)
            rc, mn, mx, err = noalloc(ctx, d_ax, N, L, d_cov, covtype, prec)
            assert rc == 0, err
            got = d_cov.get((L, L), np.float64)
            assert np.array_equal(got, ref)
            assert (mn, mx) == (mn0, mx0)
    finally:
        d_ax.free()
        d_cov.free()


def test_noalloc_rejects_oversize_and_unsupported(ctx):
    d = DevBuf(64)
    try:
        rc, _, _, err = noalloc(ctx, d, 10, 1025, d, "GTp", R.STAT_FP64)
        assert rc != 0 and "exceeds the fixed scratch" in err
        rc, _, _, err = noalloc(ctx, d, 4097, 10, d, "GTp", R.STAT_FP64)
        assert rc != 0 and "exceeds the fixed scratch" in err
        rc, _, _, err = noalloc(ctx, d, 10, 10, d, "RAFS", R.STAT_FP64)
        assert rc != 0 and "not available" in err
    finally:
        d.free()
