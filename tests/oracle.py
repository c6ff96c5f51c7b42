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


def set_threads(n):
    """Threads of the oracle's pair-count loops (results are identical for any count)."""
    lib().ref_set_threads(int(n))


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


def raf_counts(cnt, N, L, allowpair=None):
    """RAF from unweighted pair counts (closed form), as a symmetric matrix."""
    from rscape_amd import ALLOW_WCGU
    cnt = np.ascontiguousarray(cnt, dtype=np.int32)
    ap = np.ascontiguousarray(ALLOW_WCGU if allowpair is None else allowpair, dtype=np.float64)
    COV, mm = np.zeros(L * L), np.zeros(2)
    lib().ref_raf_counts(P(cnt), C.c_int64(N), C.c_int64(L), P(ap), P(COV), P(mm))
    return COV, mm


def rafs_from(B, L):
    """The RAFS stencil over a RAF matrix B (L*L, both triangles)."""
    B = np.ascontiguousarray(B, dtype=np.float64)
    COV, mm = np.zeros(L * L), np.zeros(2)
    lib().ref_rafs_from(P(B), C.c_int64(L), P(COV), P(mm))
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


class RefHist(C.Structure):
    """ref_hist of oracle/refcpu.h."""
    _fields_ = [("bmin", C.c_double), ("bmax", C.c_double), ("w", C.c_double), ("nb", C.c_int),
                ("obs", C.POINTER(C.c_uint64)), ("imin", C.c_int), ("imax", C.c_int), ("cmin", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("phi", C.c_double),
                ("n", C.c_uint64), ("Nc", C.c_uint64), ("No", C.c_uint64), ("tailfit", C.c_int)]

    def counts(self):
        return np.array([self.obs[b] for b in range(self.nb)], dtype=np.uint64)


def _hp(p):
    return C.cast(p, C.POINTER(RefHist))


def rank_hist(COV, L, maxCOV, bmin=-10.0, w=0.05, tol=1e-6):
    p = lib().ref_rank_hist(C.c_int64(L), P(np.ascontiguousarray(COV, dtype=np.float64)), maxCOV, bmin, w, tol)
    return p


def null_add(cum, h):
    cp = C.c_void_p(cum)
    lib().ref_null_add(C.byref(cp), C.c_void_p(h))
    return cp.value


def null_fit(h, pmass=5e-4, fracfit=1.0, expo=True):
    sf = C.c_void_p()
    fit = np.zeros(5)
    lib().ref_null_fit.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_void_p), C.c_void_p]
    r = lib().ref_null_fit(C.c_void_p(h), pmass, fracfit, 1 if expo else 0, C.byref(sf), P(fit))
    if r:
        raise RuntimeError("ref_null_fit failed")
    nb = _hp(h).contents.nb
    arr = None
    if sf.value:
        arr = np.ctypeslib.as_array(C.cast(sf, C.POINTER(C.c_double)), shape=(2 * nb,)).copy()
    return dict(mu=fit[0], **{"lambda": fit[1]}, tau=fit[2], newmass=fit[3], pmass=fit[4]), arr


def evalues(COV, L, Nc, h, survfit):
    Ev = np.zeros(L * L)
    lib().ref_evalues.argtypes = [C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
    lib().ref_evalues(L, P(np.ascontiguousarray(COV, dtype=np.float64)), Nc, C.c_void_p(h),
                      None if survfit is None else P(survfit), P(Ev))
    return Ev.reshape(L, L)


def evolve_gtr(tree, L, model, seed, stream, gapthr=0, root=None):
    n = tree["n"]
    out = np.zeros((n, L), dtype=np.uint8)
    f = lib().ref_evolve_gtr
    f.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    left = np.ascontiguousarray(tree["left"], dtype=np.int32); right = np.ascontiguousarray(tree["right"], dtype=np.int32)
    ld = np.ascontiguousarray(tree["ld"], dtype=np.float64); rd = np.ascontiguousarray(tree["rd"], dtype=np.float64)
    rt = None if root is None else np.ascontiguousarray(root, dtype=np.uint8)
    f(n, L, P(left), P(right), P(ld), P(rd), P(np.ascontiguousarray(model, dtype=np.float64)), seed, stream, gapthr,
      P(rt) if rt is not None else None, P(out))
    return out


def evolve_pairs(tree, L, model, m16, ct, seed, stream, gapthr=0, root=None):
    """Structured null: paired columns (ct[c] = partner or -1) evolve under the 16-state model."""
    n = tree["n"]
    out = np.zeros((n, L), dtype=np.uint8)
    f = lib().ref_evolve_pairs
    f.argtypes = [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    rt = None if root is None else np.ascontiguousarray(root, dtype=np.uint8)
    left = np.ascontiguousarray(tree["left"], dtype=np.int32); right = np.ascontiguousarray(tree["right"], dtype=np.int32)
    ld = np.ascontiguousarray(tree["ld"], dtype=np.float64); rd = np.ascontiguousarray(tree["rd"], dtype=np.float64)
    ctv = np.ascontiguousarray(ct, dtype=np.int32)
    f(n, L, P(left), P(right), P(ld), P(rd), P(np.ascontiguousarray(model, dtype=np.float64)),
      P(np.ascontiguousarray(m16, dtype=np.float64)), P(ctv), seed, stream, gapthr, P(rt) if rt is not None else None,
      P(out))
    return out


def e1_transitions(rates, time, L=100):
    """The 16 e1 HMM transitions (rows B, S, D, I of S, D, I, E) of one branch."""
    f = lib().ref_e1_transitions
    f.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.c_void_p]
    t = np.zeros(16)
    r = np.ascontiguousarray(rates, dtype=np.float64)
    if f(P(r), float(time), int(L), P(t)) != 0:
        raise RuntimeError("e1 transitions failed")
    return t


def add_indels(tree, leaves, rates, insf, seed, stream, ct=None, cap=None):
    """The indel pass over ungapped leaves (N x L0): (leaves N x alen, ct alen)."""
    n, L0 = leaves.shape
    cap = cap or 4 * L0 + 256
    f = lib().ref_add_indels
    f.argtypes = [C.c_int64, C.c_int64] + [C.c_void_p] * 6 + [C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p, C.c_int64,
                                                             C.c_void_p, C.c_void_p, C.c_void_p]
    left = np.ascontiguousarray(tree["left"], dtype=np.int32); right = np.ascontiguousarray(tree["right"], dtype=np.int32)
    ld = np.ascontiguousarray(tree["ld"], dtype=np.float64); rd = np.ascontiguousarray(tree["rd"], dtype=np.float64)
    lv = np.ascontiguousarray(leaves, dtype=np.uint8)
    ctv = None if ct is None else np.ascontiguousarray(ct, dtype=np.int32)
    while True:
        out = np.zeros((n, cap), np.uint8)
        cto = np.zeros(cap, np.int32)
        alen = C.c_int64(0)
        rc = f(n, L0, P(left), P(right), P(ld), P(rd), P(np.ascontiguousarray(rates, dtype=np.float64)),
               P(np.ascontiguousarray(insf, dtype=np.float64)), seed, stream, P(lv), P(ctv) if ctv is not None else None,
               cap, P(out), P(cto), C.byref(alen))
        if rc == 0:
            return out[:, :alen.value].copy(), cto[:alen.value].copy()
        if cap > 64 * L0 + 4096:
            raise RuntimeError("indel pass failed")
        cap *= 2


def mlp_forward_backward(W, b, x, dy=None):
    """FP64 oracle forward (and backward for dy) of the surrogate MLP."""
    nl = len(W)
    dims = [W[0].shape[1]] + [w.shape[0] for w in W]
    d = (C.c_int * (nl + 1))(*dims)
    Wd = [np.ascontiguousarray(w, dtype=np.float64) for w in W]
    bd = [np.ascontiguousarray(v, dtype=np.float64) for v in b]
    Wp = (C.c_void_p * nl)(*[w.ctypes.data for w in Wd])
    bp = (C.c_void_p * nl)(*[v.ctypes.data for v in bd])
    x = np.ascontiguousarray(x, dtype=np.float64)
    B = x.shape[0]
    aw = sum(dims[l + 1] + (dims[l + 1] if l < nl - 1 else 0) for l in range(nl))
    acts = np.zeros(B * aw)
    y = np.zeros((B, dims[-1]))
    L = lib()
    L.ref_mlp_forward(nl, d, Wp, bp, P(x), C.c_int64(B), P(y), P(acts))
    if dy is None:
        return y
    dW = [np.zeros_like(w) for w in Wd]
    db = [np.zeros_like(v) for v in bd]
    dWp = (C.c_void_p * nl)(*[w.ctypes.data for w in dW])
    dbp = (C.c_void_p * nl)(*[v.ctypes.data for v in db])
    L.ref_mlp_backward(nl, d, Wp, bp, P(x), C.c_int64(B), P(acts), P(np.ascontiguousarray(dy, dtype=np.float64)), dWp, dbp)
    return y, dW, db


def fitch_null_rows(ax, tree, seed):
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    allr = np.zeros((2 * N - 1, L), np.uint8)
    sh = np.zeros((2 * N - 1, L), np.uint8)
    left = np.ascontiguousarray(tree["left"], dtype=np.int32); right = np.ascontiguousarray(tree["right"], dtype=np.int32)
    f = lib().ref_fitch_null_rows
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
    f(P(ax), N, L, P(left), P(right), seed, P(allr), P(sh))
    return allr, sh


def write_null(path, h):
    """cov_WriteNullHistogram of an oracle histogram."""
    return lib().ref_write_null(path.encode(), h)


def read_null(path):
    """cov_ReadNullHistogram into an oracle histogram (pointer)."""
    f = lib().ref_read_null
    f.restype = C.c_void_p
    f.argtypes = [C.c_char_p]
    return f(path.encode())


def tree_substitutions(ax, tree, seed, includegaps=False):
    """Tree_Substitutions: (nsubs (L,), ndouble (L, L), njoin (L, L)) on the oracle's Fitch states."""
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    N, L = ax.shape
    ns = np.zeros(L, np.int32); nd = np.zeros(L * L, np.int32); nj = np.zeros(L * L, np.int32)
    left = np.ascontiguousarray(tree["left"], dtype=np.int32); right = np.ascontiguousarray(tree["right"], dtype=np.int32)
    f = lib().ref_tree_substitutions
    f.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int, C.c_void_p, C.c_void_p,
                  C.c_void_p]
    f(P(ax), N, L, P(left), P(right), seed, int(includegaps), P(ns), P(nd), P(nj))
    return ns, nd.reshape(L, L), nj.reshape(L, L)
