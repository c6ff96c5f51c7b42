"""ctypes binding of librscape_hip.so (the C-ABI declared in include/rscape_hip.h).

This module is a thin host-side mirror of R-scape's corr_* / cov_* call sequence
(src/covariation.c:63-306) for tests, the benchmark and scripting.  All arithmetic
runs in the HIP library; if the library is missing this module raises instead of
falling back to anything else.
"""
import ctypes as C
import os
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(os.path.dirname(_HERE), "librscape_hip.so")

OK, FAIL = 0, 1
COVTYPES = ["CHI", "CHIp", "CHIa", "GT", "GTp", "GTa", "MI", "MIp", "MIa", "MIr", "MIrp", "MIra",
            "MIg", "MIgp", "MIga", "OMES", "OMESp", "OMESa", "RAF", "RAFp", "RAFa", "RAFS", "RAFSp",
            "RAFSa", "CCF", "CCFp", "CCFa"]
COV = {n: k for k, n in enumerate(COVTYPES)}
C16, C2, CWC, CSELECT = 0, 1, 2, 3
APC, ASC = 0, 1
COUNTS_INT8, COUNTS_BF16, COUNTS_BF16X3 = 0, 1, 2
STAT_FP32, STAT_FP64 = 0, 1

# R-scape's default allowed pairs: Watson-Crick + G:U (src/R-scape.c:882-887)
ALLOW_WCGU = np.zeros(16)
for _x, _y in [(0, 3), (3, 0), (1, 2), (2, 1), (2, 3), (3, 2)]:
    ALLOW_WCGU[_x * 4 + _y] = 1.0

_lib = None


class RscapeError(RuntimeError):
    pass


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RscapeError("librscape_hip.so not built (run __graft_entry__.build())")
        _lib = C.CDLL(LIBPATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, i64, i32, d, cp = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_char_p
    sig = {
        "rsg_ctx_create": (vp, [i32, cp]),
        "rsg_ctx_destroy": (None, [vp]),
        "rsg_ctx_set_timing": (i32, [vp, i32]),
        "rsg_ctx_timings": (i32, [vp, C.POINTER(C.c_char_p), C.POINTER(d), C.POINTER(i32), i32]),
        "rsg_gram_nshapes": (i32, []),
        "rsg_gram_shape_name": (cp, [i32]),
        "rsg_ctx_set_gram_shape": (i32, [vp, i32]),
        "rsg_msa_create": (i32, [vp, vp, vp, i64, i64, C.POINTER(vp), cp]),
        "rsg_msa_update": (i32, [vp, vp, vp, vp, cp]),
        "rsg_msa_destroy": (None, [vp, vp]),
        "rsg_corr_Probs": (i32, [vp, vp, i32, d, cp]),
        "rsg_corr_GetProbs": (i32, [vp, vp, vp, vp, vp, vp, vp, cp]),
        "rsg_corr_GetCounts": (i32, [vp, vp, vp, cp]),
        "rsg_corr_CalculateCOV": (i32, [vp, vp, i32, i32, vp, i32, i32, i32, cp]),
        "rsg_corr_CalculateRAF": (i32, [vp, vp, i32, vp, i32, cp]),
        "rsg_corr_CalculateCOVCorrected": (i32, [vp, vp, i32, i32, cp]),
        "rsg_corr_GetCOV": (i32, [vp, vp, vp, C.POINTER(d), C.POINTER(d), cp]),
        "rsg_cov_Calculate": (i32, [vp, vp, i32, i32, i32, i32, vp, i32, i32, d,
                                    C.POINTER(d), C.POINTER(d), cp]),
    }
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args


def exported_symbols():
    """Names declared in include/rscape_hip.h that the library exports."""
    hdr = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "rscape_hip.h")
    import re
    names = re.findall(r"\b(rsg_[A-Za-z0-9_]+)\s*\(", open(hdr).read())
    return sorted(set(names))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _err():
    return C.create_string_buffer(128)


def _check(status, err, what):
    if status != OK:
        raise RscapeError("%s: %s" % (what, err.value.decode(errors="replace")))


class Context:
    """One device context (one HIP stream)."""

    def __init__(self, device=0):
        self.L = lib()
        err = _err()
        self.h = self.L.rsg_ctx_create(device, err)
        if not self.h:
            raise RscapeError("rsg_ctx_create: " + err.value.decode())

    def close(self):
        if self.h:
            self.L.rsg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on=True):
        self.L.rsg_ctx_set_timing(self.h, 1 if on else 0)

    def timings(self):
        names = (C.c_char_p * 64)()
        ms = (C.c_double * 64)()
        n = C.c_int(0)
        self.L.rsg_ctx_timings(self.h, names, ms, C.byref(n), 64)
        return [(names[k].decode(), ms[k]) for k in range(n.value)]

    def set_gram_shape(self, shape):
        self.L.rsg_ctx_set_gram_shape(self.h, int(shape))

    def msa(self, ax, wgt=None):
        return Msa(self, ax, wgt)


class Msa:
    """An alignment resident in HBM with its pair tables (the device side of mutual_s)."""

    def __init__(self, ctx, ax, wgt=None):
        self.ctx = ctx
        self.ax = np.ascontiguousarray(ax, dtype=np.uint8)
        self.N, self.Lc = self.ax.shape
        self.wgt = None if wgt is None else np.ascontiguousarray(wgt, dtype=np.float64)
        err = _err()
        h = C.c_void_p()
        _check(ctx.L.rsg_msa_create(ctx.h, _p(self.ax), _p(self.wgt), self.N, self.Lc, C.byref(h), err),
               err, "rsg_msa_create")
        self.h = h

    def close(self):
        if self.h:
            self.ctx.L.rsg_msa_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, ax=None, wgt=None):
        err = _err()
        a = None if ax is None else np.ascontiguousarray(ax, dtype=np.uint8)
        w = None if wgt is None else np.ascontiguousarray(wgt, dtype=np.float64)
        _check(self.ctx.L.rsg_msa_update(self.ctx.h, self.h, _p(a), _p(w), err), err, "rsg_msa_update")

    def probs(self, counts=COUNTS_INT8, tol=1e-6):
        err = _err()
        _check(self.ctx.L.rsg_corr_Probs(self.ctx.h, self.h, counts, tol, err), err, "rsg_corr_Probs")

    def counts(self):
        L = self.Lc
        out = np.zeros(L * L * 16)
        err = _err()
        _check(self.ctx.L.rsg_corr_GetCounts(self.ctx.h, self.h, _p(out), err), err, "rsg_corr_GetCounts")
        return out

    def get_probs(self):
        L = self.Lc
        pp, pm, ps = np.zeros(L * L * 16), np.zeros(L * 4), np.zeros(L * 5)
        nseff, ngap = np.zeros(L * L), np.zeros(L * L)
        err = _err()
        _check(self.ctx.L.rsg_corr_GetProbs(self.ctx.h, self.h, _p(pp), _p(pm), _p(ps), _p(nseff), _p(ngap), err),
               err, "rsg_corr_GetProbs")
        return dict(pp=pp, pm=pm, ps=ps, nseff=nseff, ngap=ngap)

    def cov(self, covtype, covclass=CSELECT, allowpair=None, nseqthresh=8, alenthresh=50, prec=STAT_FP64):
        err = _err()
        ap = None if allowpair is None else np.ascontiguousarray(allowpair, dtype=np.float64)
        base = COV[covtype] if isinstance(covtype, str) else covtype
        _check(self.ctx.L.rsg_corr_CalculateCOV(self.ctx.h, self.h, base, covclass, _p(ap), nseqthresh,
                                                alenthresh, prec, err), err, "rsg_corr_CalculateCOV")

    def correct(self, actype=APC, shiftnonneg=0):
        err = _err()
        _check(self.ctx.L.rsg_corr_CalculateCOVCorrected(self.ctx.h, self.h, actype, shiftnonneg, err),
               err, "rsg_corr_CalculateCOVCorrected")

    def get_cov(self, matrix=True):
        L = self.Lc
        out = np.zeros(L * L) if matrix else None
        mn, mx = C.c_double(), C.c_double()
        err = _err()
        _check(self.ctx.L.rsg_corr_GetCOV(self.ctx.h, self.h, _p(out), C.byref(mn), C.byref(mx), err),
               err, "rsg_corr_GetCOV")
        return (None if out is None else out.reshape(L, L)), mn.value, mx.value

    def calculate(self, covtype="GTp", covclass=CSELECT, counts=COUNTS_INT8, prec=STAT_FP64, allowpair=None,
                  nseqthresh=8, alenthresh=50, tol=1e-6):
        """cov_Calculate without the ranking: probabilities, statistic, correction."""
        err = _err()
        ap = None if allowpair is None else np.ascontiguousarray(allowpair, dtype=np.float64)
        mn, mx = C.c_double(), C.c_double()
        t = COV[covtype] if isinstance(covtype, str) else covtype
        _check(self.ctx.L.rsg_cov_Calculate(self.ctx.h, self.h, t, covclass, counts, prec, _p(ap), nseqthresh,
                                            alenthresh, tol, C.byref(mn), C.byref(mx), err), err, "rsg_cov_Calculate")
        return mn.value, mx.value
