"""Substitutions, power, hit lists and output files of a significance run (ctypes binding).

Mirrors R-scape's Tree_Substitutions (src/msatree.c:1423-1555), power_Read / power_SPAIR_Create
(src/power.c:65-217, 532-650), cov_CreateHitList (src/covariation.c:778-1006) and the .cov,
.sorted.cov, .power and .surv writers (src/covariation.c:1139-1406, 1637-1673; src/power.c:261-378).
All work happens in librscape_hip.
"""
import ctypes as C

import numpy as np

from . import Hist, _check, _err, _p, lib

SINGLE_SUBS, JOIN_SUBS, DOUBLE_SUBS = 0, 1, 2
BP_WWC, BP_STACKED, BP_CONTACT, BP_NONE = 0, 32, 33, 34
SAMPLE_CONTACTS, SAMPLE_BP, SAMPLE_WC, SAMPLE_ALL = 0, 1, 2, 3


class Power(C.Structure):
    _fields_ = [("ns", C.c_int), ("subs", C.POINTER(C.c_double)), ("prob", C.POINTER(C.c_double))]


class SPair(C.Structure):
    _fields_ = [("i", C.c_int64), ("j", C.c_int64), ("iabs", C.c_int64), ("jabs", C.c_int64),
                ("nseff", C.c_double), ("ngap", C.c_double), ("wcf_f", C.c_double), ("wobble_f", C.c_double),
                ("noncan_f", C.c_double), ("powertype", C.c_int32), ("nsubs", C.c_int64), ("nsubs_join", C.c_int64),
                ("nsubs_double", C.c_int64), ("power", C.c_double), ("power_join", C.c_double),
                ("power_double", C.c_double), ("bptype_given", C.c_int32), ("covary", C.c_int32),
                ("sc", C.c_double), ("Pval", C.c_double), ("Eval", C.c_double)]


class Hit(C.Structure):
    _fields_ = [("i", C.c_int64), ("j", C.c_int64), ("sc", C.c_double), ("Eval", C.c_double), ("pval", C.c_double),
                ("nsubs", C.c_int64), ("power", C.c_double), ("bptype", C.c_int32), ("is_compatible", C.c_int32)]


def _declare():
    L = lib()
    if getattr(L, "_output_declared", False):
        return L
    vp, i64, i32, d, cp = C.c_void_p, C.c_int64, C.c_int, C.c_double, C.c_char_p
    L.rsg_tree_Substitutions.restype = i32
    L.rsg_tree_Substitutions.argtypes = [vp, vp, i64, i64, vp, vp, C.c_uint64, i32, vp, vp, vp, cp]
    L.rsg_power_Read.restype = i32
    L.rsg_power_Read.argtypes = [cp, C.POINTER(C.POINTER(Power)), cp]
    L.rsg_power_Destroy.restype, L.rsg_power_Destroy.argtypes = None, [C.POINTER(Power)]
    L.rsg_power_Prob.restype, L.rsg_power_Prob.argtypes = d, [C.POINTER(Power), i64]
    L.rsg_power_SPAIR_Create.restype = i32
    L.rsg_power_SPAIR_Create.argtypes = [vp, vp, vp, C.POINTER(Power), i32, vp, vp, vp, vp, C.POINTER(SPair), cp]
    L.rsg_cov_CreateHitList.restype = i32
    L.rsg_cov_CreateHitList.argtypes = [vp, vp, C.POINTER(Hist), C.POINTER(d), i32, vp, vp, i64, i64, d, d,
                                        C.POINTER(SPair), C.POINTER(C.POINTER(Hit)), C.POINTER(i32), cp]
    L.rsg_cov_WriteHitList.restype = i32
    L.rsg_cov_WriteHitList.argtypes = [cp, i32, i32, cp, d, d, d, i32, i64, i64, C.POINTER(Hit), i32, C.POINTER(SPair),
                                       i64, vp, i32, i32, cp]
    L.rsg_power_SPAIR_Write.restype = i32
    L.rsg_power_SPAIR_Write.argtypes = [cp, C.POINTER(SPair), i64, i32, cp]
    L.rsg_cov_WriteSurv.restype = i32
    L.rsg_cov_WriteSurv.argtypes = [cp, C.POINTER(Hist), C.POINTER(Hist), C.POINTER(Hist), C.POINTER(d), cp]
    L._output_declared = True
    return L


def substitutions(ctx, ax, tree, seed=1, includegaps=False, pairs=True):
    """(nsubs (L,), ndouble (L, L), njoin (L, L)); the pair arrays are None without `pairs`."""
    L_ = _declare()
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    left = np.ascontiguousarray(tree["left"], dtype=np.int32)
    right = np.ascontiguousarray(tree["right"], dtype=np.int32)
    ns = np.zeros(L, np.int32)
    nd = np.zeros((L, L), np.int32) if pairs else None
    nj = np.zeros((L, L), np.int32) if pairs else None
    err = _err()
    _check(L_.rsg_tree_Substitutions(ctx.h, _p(ax), N, L, _p(left), _p(right), seed, int(includegaps), _p(ns), _p(nd),
                                     _p(nj), err), err, "rsg_tree_Substitutions")
    return ns, nd, nj


class PowerTable:
    def __init__(self, path):
        L_ = _declare()
        self.p = C.POINTER(Power)()
        err = _err()
        _check(L_.rsg_power_Read(path.encode(), C.byref(self.p), err), err, "rsg_power_Read")

    def prob(self, nsubs):
        return _declare().rsg_power_Prob(self.p, int(nsubs))

    def table(self):
        n = self.p.contents.ns
        return (np.ctypeslib.as_array(self.p.contents.subs, shape=(n,)).copy(),
                np.ctypeslib.as_array(self.p.contents.prob, shape=(n,)).copy()) if n else (np.zeros(0), np.zeros(0))

    def __del__(self):
        try:
            if self.p:
                _declare().rsg_power_Destroy(self.p)
                self.p = None
        except Exception:
            pass


def spairs(ctx, msa, power=None, ct=None, nsubs=None, njoin=None, ndouble=None, msamap=None):
    """power_SPAIR_Create + AddMI after corr_Probs: a ctypes array of L (L - 1) / 2 SPair."""
    L_ = _declare()
    L = msa.Lc
    dim = L * (L - 1) // 2
    arr = (SPair * max(dim, 1))()
    conv = lambda a, t: None if a is None else np.ascontiguousarray(a, dtype=t)
    ctv, nsv, njv, ndv, mm = conv(ct, np.int32), conv(nsubs, np.int32), conv(njoin, np.int32), conv(ndouble, np.int32), \
        conv(msamap, np.int32)
    err = _err()
    _check(L_.rsg_power_SPAIR_Create(ctx.h, msa.h, _p(mm), power.p if power is not None else None, 0, _p(ctv), _p(nsv),
                                     _p(njv), _p(ndv), arr, err), err, "rsg_power_SPAIR_Create")
    arr._keep = (ctv, nsv, njv, ndv, mm)
    return arr


def hit_list(ctx, msa, null, survfit, samplesize=SAMPLE_ALL, ct=None, bptype=None, Nt=None, Nb=0, expBP=-1.0,
             thresh=0.05, spair=None):
    """cov_CreateHitList's selection: a list of Hit structures (copies)."""
    L_ = _declare()
    L = msa.Lc
    Nt = L * (L - 1) // 2 if Nt is None else Nt
    ctv = None if ct is None else np.ascontiguousarray(ct, dtype=np.int32)
    bpv = None if bptype is None else np.ascontiguousarray(bptype, dtype=np.uint8)
    sfp = None if survfit is None else survfit.ctypes.data_as(C.POINTER(C.c_double))
    hp = C.POINTER(Hit)()
    n = C.c_int(0)
    err = _err()
    _check(L_.rsg_cov_CreateHitList(ctx.h, msa.h, null.p if null is not None else None, sfp, samplesize, _p(ctv), _p(bpv),
                                    int(Nt), int(Nb), float(expBP), float(thresh), spair, C.byref(hp), C.byref(n), err),
           err, "rsg_cov_CreateHitList")
    hits = [Hit.from_buffer_copy(hp[k]) for k in range(n.value)]
    lib().rsg_free(C.cast(hp, C.c_void_p))
    return hits


def write_hits(path, hits, spair, alen, covtype="GTp", thresh=0.05, cov_min=0.0, cov_max=0.0, tnum=0, Nb=0, Nt=0,
               ranked=False, by_score=False, msamap=None, firstpos=1, samplesize=SAMPLE_ALL):
    L_ = _declare()
    arr = (Hit * max(len(hits), 1))(*hits)
    mm = None if msamap is None else np.ascontiguousarray(msamap, dtype=np.int32)
    err = _err()
    _check(L_.rsg_cov_WriteHitList(path.encode(), int(ranked), int(by_score), covtype.encode(), thresh, cov_min, cov_max,
                                   int(tnum), int(Nb), int(Nt), arr, len(hits), spair, int(alen), _p(mm), int(firstpos),
                                   int(samplesize), err), err, "rsg_cov_WriteHitList")


def write_power(path, spair, dim, in_given=True):
    L_ = _declare()
    err = _err()
    _check(L_.rsg_power_SPAIR_Write(path.encode(), spair, int(dim), int(in_given), err), err, "rsg_power_SPAIR_Write")


def write_surv(path, hb=None, ht=None, null=None, survfit=None):
    L_ = _declare()
    err = _err()
    sfp = None if survfit is None else survfit.ctypes.data_as(C.POINTER(C.c_double))
    g = lambda h: None if h is None else h.p
    _check(L_.rsg_cov_WriteSurv(path.encode(), g(hb), g(ht), g(null), sfp, err), err, "rsg_cov_WriteSurv")
