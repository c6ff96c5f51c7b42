"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None

CHI, GT, MI, MIr, MIg, OMES, RAF, RAFS, CCF = 0, 3, 6, 9, 12, 15, 18, 21, 24
C16, C2, CWC, CSELECT = 0, 1, 2, 3
APC, ASC = 0, 1


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        _lib = C.CDLL(LIB)
        d, i64, vp = C.c_double, C.c_int64, C.c_void_p
        _lib.ref_cov2evalue.restype = d
        _lib.ref_evalue2cov.restype = d
        _lib.ref_exp_surv.restype = d
        _lib.ref_gam_surv.restype = d
        _lib.ref_hist_pmass.restype = d
        _lib.ref_width.restype = d
        _lib.ref_width.argtypes = [d, d, d, d, C.c_int, d]
        _lib.ref_hist_create.restype = vp
        _lib.ref_hist_create.argtypes = [d, d, d]
        _lib.ref_rank_hist.restype = vp
        _lib.ref_rank_hist.argtypes = [i64, vp, d, d, d, d]
        _lib.ref_survfit.restype = vp
        _lib.ref_rng_u32.restype = C.c_uint32
        _lib.ref_rng_u32.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
    return _lib


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def probs(ax, wgt=None, tol=1e-6):
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    w = np.ones(N) if wgt is None else np.ascontiguousarray(wgt, dtype=np.float64)
    pp, nseff, ngap = np.zeros(L * L * 16), np.zeros(L * L), np.zeros(L * L)
    pm, ps = np.zeros(L * 4), np.zeros(L * 5)
    err = C.create_string_buffer(256)
    r = lib().ref_probs(P(ax), P(w), C.c_int64(N), C.c_int64(L), C.c_double(tol), P(pp), P(nseff), P(ngap),
                        P(pm), P(ps), err)
    if r:
        raise RuntimeError(err.value)
    return dict(pp=pp, nseff=nseff, ngap=ngap, pm=pm, ps=ps, N=N, L=L)


def pair_counts(ax):
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    cnt = np.zeros(L * L * 16, dtype=np.int32)
    lib().ref_pair_counts(P(ax), C.c_int64(N), C.c_int64(L), P(cnt))
    return cnt


def cov(pr, covtype, covclass=CSELECT, allowpair=None, nseqthresh=8, alenthresh=50):
    from rscape_amd import ALLOW_WCGU
    L = pr["L"]
    ap = np.ascontiguousarray(ALLOW_WCGU if allowpair is None else allowpair, dtype=np.float64)
    COV, mm = np.zeros(L * L), np.zeros(2)
    err = C.create_string_buffer(256)
    r = lib().ref_cov_stat(covtype, covclass, C.c_int64(pr["N"]), C.c_int64(L), nseqthresh, alenthresh, P(ap),
                           P(pr["pp"]), P(pr["pm"]), P(pr["nseff"]), P(pr["ngap"]), P(COV), P(mm), err)
    if r:
        raise RuntimeError(err.value)
    return COV, mm


def raf(ax, rafs=False, allowpair=None):
    from rscape_amd import ALLOW_WCGU
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    ap = np.ascontiguousarray(ALLOW_WCGU if allowpair is None else allowpair, dtype=np.float64)
    COV, mm = np.zeros(L * L), np.zeros(2)
    f = lib().ref_rafs if rafs else lib().ref_raf
    f(P(ax), C.c_int64(N), C.c_int64(L), P(ap), P(COV), P(mm))
    return COV, mm


def corrected(COV, L, actype=APC, shiftnonneg=0):
    COV = COV.copy()
    mm = np.zeros(2)
    err = C.create_string_buffer(256)
    r = lib().ref_cov_corrected(actype, C.c_int64(L), P(COV), P(mm), shiftnonneg, err)
    if r:
        raise RuntimeError(err.value)
    return COV, mm


def weights_gsc(ax):
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    w = np.zeros(N)
    lib().ref_weights_gsc(P(ax), C.c_int64(N), C.c_int64(L), P(w))
    return w


def weights_pb(ax):
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    w = np.zeros(N)
    lib().ref_weights_pb(P(ax), C.c_int64(N), C.c_int64(L), P(w))
    return w


def gapcol_filter(ax, wgt, gapthresh=0.75):
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    use = np.zeros(L, dtype=np.int32)
    lib().ref_gapcol_filter(P(ax), P(np.ascontiguousarray(wgt, dtype=np.float64)), C.c_int64(N), C.c_int64(L),
                            C.c_double(gapthresh), P(use))
    return use.astype(bool)


def canonicalize(ax):
    ax = np.ascontiguousarray(ax, dtype=np.uint8).copy()
    lib().ref_canonicalize(P(ax), C.c_int64(ax.size))
    return ax
