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
COUNTS_INT8, COUNTS_BF16, COUNTS_BF16X3, COUNTS_W8 = 0, 1, 2, 3
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
        "rsg_gram_tiles_load": (i32, [cp, cp]),
        "rsg_ctx_set_gram_split": (i32, [vp, i32]),
        "rsg_ctx_set_batch_persistent": (i32, [vp, i32]),
        "rsg_ctx_set_stamp_wait": (i32, [vp, i32]),
        "rsg_ctx_graph_reuses": (i32, [vp]),
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
    hp = C.POINTER(Hist)
    sig.update({
        "rsg_width": (d, [d, d, d, d, i32, d]),
        "rsg_cov_RankHist": (i32, [vp, vp, d, d, d, C.POINTER(hp), cp]),
        "rsg_hist_NullAdd": (i32, [C.POINTER(hp), hp, cp]),
        "rsg_hist_NullFit": (i32, [vp, hp, d, d, i32, C.POINTER(C.POINTER(d)), C.POINTER(d), cp]),
        "rsg_cov_Evalues": (i32, [vp, vp, hp, C.POINTER(d), i64, vp, cp]),
        "rsg_cov2evalue": (d, [d, i64, hp, C.POINTER(d)]),
        "rsg_evalue2cov": (d, [d, i64, hp, C.POINTER(d)]),
        "rsg_hist_create": (hp, [d, d, d]),
        "rsg_hist_destroy": (None, [hp]),
        "rsg_free": (None, [vp]),
        "rsg_hist_WriteNull": (i32, [cp, hp, cp]),
        "rsg_hist_ReadNull": (i32, [cp, C.POINTER(hp), cp]),
        "rsg_weights_PB": (i32, [vp, vp, i64, i64, vp, cp]),
        "rsg_weights_GSC": (i32, [vp, vp, i64, i64, vp, cp]),
        "rsg_gapcol_mask": (i32, [vp, vp, i64, i64, d, vp]),
        "rsg_msa_RemoveGapColumns": (i32, [vp, vp, d, vp, vp, C.POINTER(vp), cp]),
    })
    for name, (res, args) in sig.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args


class Hist(C.Structure):
    """rsg_hist (include/rscape_hip.h): Easel-style score histogram."""
    _fields_ = [("bmin", C.c_double), ("bmax", C.c_double), ("w", C.c_double),
                ("nb", C.c_int32), ("imin", C.c_int32), ("imax", C.c_int32), ("cmin", C.c_int32),
                ("xmin", C.c_double), ("xmax", C.c_double), ("phi", C.c_double),
                ("n", C.c_uint64), ("Nc", C.c_uint64), ("No", C.c_uint64), ("tailfit", C.c_int32),
                ("obs", C.POINTER(C.c_uint64))]

    def counts(self):
        if self.nb <= 0:
            return np.zeros(0, dtype=np.uint64)
        return np.ctypeslib.as_array(self.obs, shape=(self.nb,)).astype(np.uint64, copy=True)


class Histogram:
    """Owner of an rsg_hist allocated by the library."""

    def __init__(self, ptr):
        self.p = ptr

    @property
    def h(self):
        return self.p.contents

    def __del__(self):
        try:
            if self.p:
                lib().rsg_hist_destroy(self.p)
                self.p = None
        except Exception:
            pass

    def null_add(self_or_none, other):
        """Merge `other` into this null histogram (null_add2cumranklist); returns the merged one."""
        err = _err()
        p = C.POINTER(Hist)() if self_or_none is None else self_or_none.p
        _check(lib().rsg_hist_NullAdd(C.byref(p), other.p, err), err, "rsg_hist_NullAdd")
        if self_or_none is None:
            return Histogram(p)
        self_or_none.p = p
        return self_or_none

    def fit(self, ctx, pmass=5e-4, fracfit=1.0, expo=True):
        """Tail fit on the device; returns (fit dict, survfit array or None)."""
        err = _err()
        sf = C.POINTER(C.c_double)()
        out = (C.c_double * 5)()
        _check(lib().rsg_hist_NullFit(ctx.h, self.p, pmass, fracfit, 1 if expo else 0, C.byref(sf), out, err),
               err, "rsg_hist_NullFit")
        fit = dict(mu=out[0], **{"lambda": out[1]}, tau=out[2], newmass=out[3], pmass=out[4])
        if sf:
            arr = np.ctypeslib.as_array(sf, shape=(2 * self.h.nb,)).copy()
            lib().rsg_free(C.cast(sf, C.c_void_p))
            return fit, arr
        return fit, None

    def cov2evalue(self, cov, Nc, survfit=None):
        sfp = None if survfit is None else survfit.ctypes.data_as(C.POINTER(C.c_double))
        return lib().rsg_cov2evalue(cov, Nc, self.p, sfp)

    def evalue2cov(self, E, Nc, survfit=None):
        sfp = None if survfit is None else survfit.ctypes.data_as(C.POINTER(C.c_double))
        return lib().rsg_evalue2cov(E, Nc, self.p, sfp)

    def write_null(self, path):
        err = _err()
        _check(lib().rsg_hist_WriteNull(path.encode(), self.p, err), err, "rsg_hist_WriteNull")

    @staticmethod
    def read_null(path):
        err = _err()
        p = C.POINTER(Hist)()
        _check(lib().rsg_hist_ReadNull(path.encode(), C.byref(p), err), err, "rsg_hist_ReadNull")
        return Histogram(p)


def width(maxCOV, minCOV, bmin=-10.0, w_old=0.05, hpts=400, tol=1e-6):
    """calculate_width_histo's histogram step (R-scape.c:1353-1360)."""
    return lib().rsg_width(maxCOV, minCOV, bmin, w_old, hpts, tol)


def gapcol_mask(ax, wgt=None, gapthresh=0.75):
    """msamanip_RemoveGapColumns' column test: (useme bool array, kept count)."""
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    useme = np.zeros(ax.shape[1], dtype=np.int32)
    w = None if wgt is None else np.ascontiguousarray(wgt, dtype=np.float64)
    kept = lib().rsg_gapcol_mask(_p(ax), _p(w), ax.shape[0], ax.shape[1], gapthresh, _p(useme))
    if kept < 0:
        raise RscapeError("rsg_gapcol_mask: bad arguments")
    return useme.astype(bool), kept


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
        """True: time every stage; "gram": the Gram launches only (two events per scoring)."""
        self.L.rsg_ctx_set_timing(self.h, 2 if on == "gram" else (1 if on else 0))

    def timings(self):
        names = (C.c_char_p * 64)()
        ms = (C.c_double * 64)()
        n = C.c_int(0)
        self.L.rsg_ctx_timings(self.h, names, ms, C.byref(n), 64)
        return [(names[k].decode(), ms[k]) for k in range(n.value)]

    def set_gram_shape(self, shape):
        self.L.rsg_ctx_set_gram_shape(self.h, int(shape))

    def set_gram_split(self, on):
        self.L.rsg_ctx_set_gram_split(self.h, 1 if on else 0)

    def set_batch_persistent(self, on):
        """Batched null: persistent kernels (default) or one work group per work unit."""
        self.L.rsg_ctx_set_batch_persistent(self.h, 1 if on else 0)

    def set_stamp_wait(self, microseconds):
        """How long a scoring call spins on its completion stamp before it synchronises the
        stream instead (0: always synchronise)."""
        self.L.rsg_ctx_set_stamp_wait(self.h, int(microseconds))

    def graph_reuses(self):
        """Batched-null calls that replayed the context's cached graph without a new capture."""
        return int(self.L.rsg_ctx_graph_reuses(self.h))

    def msa(self, ax, wgt=None):
        return Msa(self, ax, wgt)

    def weights(self, ax, method="PB"):
        """esl_msaweight_PB / esl_msaweight_GSC weights of an N x L code array (FP64, sum N)."""
        ax = np.ascontiguousarray(ax, dtype=np.uint8)
        w = np.zeros(ax.shape[0])
        err = _err()
        f = {"PB": self.L.rsg_weights_PB, "GSC": self.L.rsg_weights_GSC}[method]
        _check(f(self.h, _p(ax), ax.shape[0], ax.shape[1], _p(w), err), err, "rsg_weights_" + method)
        return w

    def msaweight(self, ax, maxsq_gsc=1000):
        """R-scape's choice (R-scape.c:1554-1557): GSC up to maxsq_gsc sequences, PB above."""
        return self.weights(ax, "GSC" if np.asarray(ax).shape[0] <= maxsq_gsc else "PB")


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

    @classmethod
    def _wrap(cls, ctx, h, N, L, wgt):
        m = cls.__new__(cls)
        m.ctx, m.h, m.N, m.Lc, m.wgt, m.ax = ctx, h, N, L, wgt, None
        return m

    def remove_gap_columns(self, gapthresh):
        """msamanip_RemoveGapColumns on the device: (resident alignment of the kept columns,
        useme (L,), map (kept,) new -> old column)."""
        useme = np.zeros(self.Lc, np.int32)
        mp = np.zeros(self.Lc, np.int32)
        h = C.c_void_p()
        err = _err()
        _check(self.ctx.L.rsg_msa_RemoveGapColumns(self.ctx.h, self.h, float(gapthresh), _p(useme), _p(mp), C.byref(h),
                                                   err), err, "rsg_msa_RemoveGapColumns")
        k = int(useme.sum())
        return Msa._wrap(self.ctx, h, self.N, k, self.wgt), useme, mp[:k].copy()

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

    def rank_hist(self, bmin=-10.0, w=0.05, tol=1e-6):
        """Histogram of max(COV_ij, bmin + w) over i<j (cov_SignificantPairs_Ranking)."""
        err = _err()
        p = C.POINTER(Hist)()
        _check(self.ctx.L.rsg_cov_RankHist(self.ctx.h, self.h, bmin, w, tol, C.byref(p), err), err, "rsg_cov_RankHist")
        return Histogram(p)

    def evalues(self, null, survfit, Nc=None, matrix=True):
        """E-value of every pair against a fitted null (cov2evalue * Nc); with matrix=False the
        values stay on the device (for null.hits())."""
        L = self.Lc
        Nc = L * (L - 1) // 2 if Nc is None else Nc
        out = np.zeros(L * L) if matrix else None
        err = _err()
        sfp = None if survfit is None else survfit.ctypes.data_as(C.POINTER(C.c_double))
        _check(self.ctx.L.rsg_cov_Evalues(self.ctx.h, self.h, null.p, sfp, Nc, _p(out), err), err, "rsg_cov_Evalues")
        return None if out is None else out.reshape(L, L)

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
