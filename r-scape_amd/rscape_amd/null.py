"""Phylogenetic null simulation and the batched null pipeline (ctypes binding).

Trees follow the Easel ESL_TREE layout (internal nodes 0..N-2, root 0, parents first;
children > 0 internal, <= 0 leaf -child).  Substitution model: 4-state GTR scaled to one
expected substitution per unit branch length.  All arithmetic runs in librscape_hip.
"""
import ctypes as C
import math

import numpy as np

from . import _check, _err, _p, lib, COV

# R-scape-sim-like defaults: uniform composition, transitions twice transversions
UNIFORM = (0.25, 0.25, 0.25, 0.25)
EXCH = (1.0, 2.0, 1.0, 1.0, 2.0, 1.0)          # AC AG AU CG CU GU


class NullJob(C.Structure):
    _fields_ = [("nleaf", C.c_int64), ("alen", C.c_int64),
                ("left", C.POINTER(C.c_int)), ("right", C.POINTER(C.c_int)),
                ("ld", C.POINTER(C.c_double)), ("rd", C.POINTER(C.c_double)),
                ("model", C.POINTER(C.c_double)), ("seed", C.c_uint64), ("stream", C.c_uint32),
                ("gapfrac", C.c_double), ("group", C.c_int32)]


_JOB_DTYPE = np.dtype({"names": [f[0] for f in NullJob._fields_],
                       "formats": ["<i8", "<i8", "<u8", "<u8", "<u8", "<u8", "<u8", "<u8", "<u4", "<f8", "<i4"],
                       "offsets": [getattr(NullJob, f[0]).offset for f in NullJob._fields_],
                       "itemsize": C.sizeof(NullJob)})


def _addr(a):
    return a.__array_interface__["data"][0]


def job_array(jobs):
    """The jobs as one rsg_null_job array (numpy records in the C layout; the jobs' arrays
    must stay alive while it is used).  Jobs with group=None get group = their index.  One
    tuple per job (the array addresses are taken once, when the job is made)."""
    rows = [(j.tree["n"], j.L) + j.addrs() + (j.seed, j.stream, j.gapfrac, k if j.group is None else j.group)
            for k, j in enumerate(jobs)]
    return np.array(rows, dtype=_JOB_DTYPE) if rows else np.zeros(0, _JOB_DTYPE)


class E1Rates(C.Structure):
    """rsg_e1_rates: the E1_RATE fields of the e1 indel model's transitions."""
    _fields_ = [(k, C.c_double) for k in ("rI", "rM", "rD", "muAM", "muAD", "muAI", "ldI", "muI")]


def aif_rates(rI, muAM, ldI, muI):
    """The AIF model's rates (src/e1_rate.c:431-519): rM = rD = 0, muAD = muAI = muAM."""
    return E1Rates(rI, 0.0, 0.0, muAM, muAM, muAM, ldI, muI)


# R-scape-sim's default evomodel parameters (data/training/Pfam.seed.S1000.trainGD.AIF.param:
# rI 0.846315, ldI 0.0101, muI 0.306055, muAM 0.0102)
AIF_PFAM = (0.846315, 0.0102, 0.0101, 0.306055)


def _declare():
    L = lib()
    if getattr(L, "_null_declared", False):
        return L
    d, i64, vp, cp, i32 = C.c_double, C.c_int64, C.c_void_p, C.c_char_p, C.c_int
    L.rsg_gtr_model.restype, L.rsg_gtr_model.argtypes = i32, [vp, vp, vp]
    L.rsg_gtr_cdf.restype, L.rsg_gtr_cdf.argtypes = i32, [vp, vp, i64, vp, vp, cp]
    L.rsg_tree_yule.restype = i32
    L.rsg_tree_yule.argtypes = [i64, d, C.c_uint64, C.c_uint32, vp, vp, vp, vp]
    L.rsg_null_simulate.restype, L.rsg_null_simulate.argtypes = i32, [vp, C.POINTER(NullJob), vp, cp]
    L.rsg_pair_model.restype, L.rsg_pair_model.argtypes = i32, [vp, vp, vp]
    L.rsg_null_simulate_pairs.restype = i32
    L.rsg_null_simulate_pairs.argtypes = [vp, C.POINTER(NullJob), vp, vp, vp, cp]
    L.rsg_null_fitch.restype = i32
    L.rsg_null_fitch.argtypes = [vp, vp, i64, i64, vp, vp, C.c_uint64, vp, cp]
    L.rsg_null_batch.restype = i32
    L.rsg_null_batch.argtypes = [vp, vp, i32, i32, d, d, d, i32, vp, vp, vp, i32, cp]
    L.rsg_null_batch_fitch.restype = i32
    L.rsg_null_batch_fitch.argtypes = [vp, vp, i64, i64, vp, vp, vp, i32, i32, d, d, d, i32, vp, vp, i32, cp]
    L.rsg_e1_transitions.restype = i32
    L.rsg_e1_transitions.argtypes = [C.POINTER(E1Rates), d, i64, vp]
    L.rsg_cov_GenerateAlignment.restype = i32
    L.rsg_cov_GenerateAlignment.argtypes = [vp, C.POINTER(NullJob), vp, vp, vp, C.POINTER(E1Rates), vp,
                                            C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(i64), C.POINTER(C.POINTER(C.c_int)), cp]
    L.rsg_null_features.restype = i32
    L.rsg_null_features.argtypes = [vp, vp, vp, vp, vp, vp, vp, cp]
    L.rsg_cov_EvaluesTail.restype = i32
    L.rsg_cov_EvaluesTail.argtypes = [vp, vp, d, d, d, d, i64, vp, cp]
    L.rsg_cov_Hits.restype = i32
    L.rsg_cov_Hits.argtypes = [vp, vp, d, i32, C.POINTER(i32), vp, vp, vp, vp, cp]
    L._null_declared = True
    return L


def gtr_model(pi=UNIFORM, exch=EXCH):
    """Eigendecomposed GTR model (40 doubles) for the device."""
    L = _declare()
    m = np.zeros(40)
    p = np.ascontiguousarray(pi, dtype=np.float64)
    e = np.ascontiguousarray(exch, dtype=np.float64)
    if L.rsg_gtr_model(_p(p), _p(e), _p(m)) != 0:
        raise ValueError("bad GTR parameters")
    return m


def gtr_cdf(ctx, model, t):
    """Device-built inverse-CDF thresholds of P(t) for each t: array (len(t), 4, 3) uint32."""
    L = _declare()
    t = np.ascontiguousarray(t, dtype=np.float64)
    out = np.zeros(len(t) * 12, dtype=np.uint32)
    err = _err()
    _check(L.rsg_gtr_cdf(ctx.h, _p(model), len(t), _p(t), _p(out), err), err, "rsg_gtr_cdf")
    return out.reshape(len(t), 4, 3)


def yule_tree(n, mean_bl, seed=42, stream=0):
    """Yule tree (dict with left, right, ld, rd, n) from the library's generator."""
    L = _declare()
    left = np.zeros(n - 1, dtype=np.int32); right = np.zeros(n - 1, dtype=np.int32)
    ld = np.zeros(n - 1); rd = np.zeros(n - 1)
    if L.rsg_tree_yule(n, mean_bl, seed, stream, _p(left), _p(right), _p(ld), _p(rd)) != 0:
        raise ValueError("tree needs n >= 2")
    return dict(left=left, right=right, ld=ld, rd=rd, n=n)


def star_tree(n, bl):
    """R-scape-sim's star tree (cov_evolve_root_ungapped_star, src/cov_simulate.c:241-262: every
    leaf evolves from the root over the same time bl) as a binary tree the simulator walks by
    depth: a complete binary tree in heap order (internal nodes 0..n-2, node v's children
    2v+1 and 2v+2, leaves numbered in order), internal branches of length 0 (P(0) = I, no
    change), leaf branches of length bl.  Depth ceil(log2 n) instead of n - 1 for a comb."""
    if n < 2:
        raise ValueError("a tree needs n >= 2")
    m = n - 1
    left = np.zeros(m, dtype=np.int32); right = np.zeros(m, dtype=np.int32)
    ld = np.zeros(m); rd = np.zeros(m)
    for v in range(m):
        for side, c in ((0, 2 * v + 1), (1, 2 * v + 2)):
            d, t = (c, 0.0) if c < m else (-(c - m), float(bl))
            (left if side == 0 else right)[v] = d
            (ld if side == 0 else rd)[v] = t
    return dict(left=left, right=right, ld=ld, rd=rd, n=n)


def rand_tree(n):
    """R-scape-sim's random alignments (cov_evolve_root_ungapped_rand, src/cov_simulate.c:265-285:
    every residue drawn from the background frequencies, independently): a star tree whose leaf
    branches are long enough that P(t) is the stationary distribution to double precision."""
    return star_tree(n, 1e3)


class Job:
    """One alignment to simulate: tree, length, model and RNG key."""

    def __init__(self, tree, L, model, seed=42, stream=0, gapfrac=0.0, group=None):
        self.tree = {k: (np.ascontiguousarray(v, dtype=np.int32) if k in ("left", "right") else
                         np.ascontiguousarray(v, dtype=np.float64) if k in ("ld", "rd") else v)
                     for k, v in tree.items()}
        self.L, self.model = int(L), np.ascontiguousarray(model, dtype=np.float64)
        self.seed, self.stream, self.gapfrac, self.group = seed, stream, gapfrac, group
        self._addrs = None

    def addrs(self):
        """Addresses of the tree arrays (left, right, ld, rd) and the model, for the C records."""
        if self._addrs is None:
            t = self.tree
            self._addrs = (_addr(t["left"]), _addr(t["right"]), _addr(t["ld"]), _addr(t["rd"]), _addr(self.model))
        return self._addrs

    def struct(self, group=0):
        t = self.tree
        ip = C.POINTER(C.c_int); dp = C.POINTER(C.c_double)
        return NullJob(t["n"], self.L, t["left"].ctypes.data_as(ip), t["right"].ctypes.data_as(ip),
                       t["ld"].ctypes.data_as(dp), t["rd"].ctypes.data_as(dp), self.model.ctypes.data_as(dp),
                       self.seed, self.stream, self.gapfrac, group if self.group is None else self.group)


def simulate(ctx, job):
    """Leaves (N x L uint8) of one simulated alignment."""
    L = _declare()
    out = np.zeros((job.tree["n"], job.L), dtype=np.uint8)
    s = job.struct()
    err = _err()
    _check(L.rsg_null_simulate(ctx.h, C.byref(s), _p(out), err), err, "rsg_null_simulate")
    return out


def pair_model(Q16, M16):
    """Eigendecomposed 16-state pair model (544 doubles) from a reversible pair rate matrix and
    its stationary pair frequencies (states a*4 + b)."""
    L = _declare()
    q = np.ascontiguousarray(Q16, dtype=np.float64).reshape(256)
    m = np.ascontiguousarray(M16, dtype=np.float64).reshape(16)
    out = np.zeros(544)
    if L.rsg_pair_model(_p(q), _p(m), _p(out)) != 0:
        raise ValueError("bad or non-reversible pair model")
    return out


def simulate_pairs(ctx, job, ct, model16):
    """Structured simulation: columns paired by ct (partner or -1) evolve under model16."""
    L = _declare()
    s = job.struct(0)
    out = np.zeros((job.tree["n"], job.L), dtype=np.uint8)
    ctv = np.ascontiguousarray(ct, dtype=np.int32)
    m16 = np.ascontiguousarray(model16, dtype=np.float64)
    err = _err()
    _check(L.rsg_null_simulate_pairs(ctx.h, C.byref(s), _p(ctv), _p(m16), _p(out), err), err, "rsg_null_simulate_pairs")
    return out


def e1_transitions(rates, time, L=100):
    """The 16 transitions (rows B, S, D, I of S, D, I, E) of a branch of length time."""
    t = np.zeros(16)
    if _declare().rsg_e1_transitions(C.byref(rates), float(time), int(L), _p(t)) != 0:
        raise RuntimeError("e1 transitions failed")
    return t


def generate_alignment(ctx, job, root=None, ct=None, model16=None, rates=None, insf=None):
    """cov_GenerateAlignment on the device: (leaves N x alen, ct alen).  rates None: no indels;
    ct and model16 given: base pairs evolve under the 16-state model."""
    L = _declare()
    s = job.struct()
    rt = None if root is None else np.ascontiguousarray(root, dtype=np.uint8)
    ctv = None if ct is None else np.ascontiguousarray(ct, dtype=np.int32)
    m16 = None if model16 is None else np.ascontiguousarray(model16, dtype=np.float64)
    f = None if insf is None else np.ascontiguousarray(insf, dtype=np.float64)
    out = C.POINTER(C.c_uint8)()
    cto = C.POINTER(C.c_int)()
    alen = C.c_int64(0)
    err = _err()
    _check(L.rsg_cov_GenerateAlignment(ctx.h, C.byref(s), _p(rt), _p(ctv), _p(m16), C.byref(rates) if rates is not None else None,
                                       _p(f), C.byref(out), C.byref(alen), C.byref(cto), err), err, "rsg_cov_GenerateAlignment")
    n, a = job.tree["n"], alen.value
    leaves = np.ctypeslib.as_array(out, shape=(n, a)).copy()
    ctn = np.ctypeslib.as_array(cto, shape=(a,)).copy()
    lib().rsg_free(C.cast(out, C.c_void_p)); lib().rsg_free(C.cast(cto, C.c_void_p))
    return leaves, ctn


def fitch_null(ctx, ax, tree, seed=1):
    """R-scape's null alignment of ax on its tree (Fitch ancestors, column shuffle,
    substitution replay): N x L codes 0..4."""
    L = _declare()
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    left = np.ascontiguousarray(tree["left"], dtype=np.int32); right = np.ascontiguousarray(tree["right"], dtype=np.int32)
    out = np.zeros_like(ax)
    err = _err()
    _check(L.rsg_null_fitch(ctx.h, _p(ax), ax.shape[0], ax.shape[1], _p(left), _p(right), seed, _p(out), err),
           err, "rsg_null_fitch")
    return out


def batch_fitch(ctx, ax, tree, seeds, covtype="GTp", bmin=-10.0, pmass=5e-4, fracfit=1.0, expo=True, graph=0):
    """R-scape's null of one alignment (null_rscape's nshuffle loop) in one device batch: one
    Fitch + shuffle replicate per seed, each scored, all merged into one histogram and one
    tail fit.  Returns (fit (8,), histogram)."""
    from . import Hist, Histogram
    L = _declare()
    ax = np.ascontiguousarray(ax, dtype=np.uint8)
    left = np.ascontiguousarray(tree["left"], dtype=np.int32); right = np.ascontiguousarray(tree["right"], dtype=np.int32)
    sd = np.ascontiguousarray(seeds, dtype=np.uint64)
    fit = np.zeros(8)
    hp = (C.POINTER(Hist) * 1)()
    err = _err()
    t = COV[covtype] if isinstance(covtype, str) else covtype
    _check(L.rsg_null_batch_fitch(ctx.h, _p(ax), ax.shape[0], ax.shape[1], _p(left), _p(right), _p(sd), len(sd), t,
                                  bmin, pmass, fracfit, 1 if expo else 0, C.cast(hp, C.c_void_p), _p(fit), int(graph),
                                  err), err, "rsg_null_batch_fitch")
    return fit, (Histogram(hp[0]) if hp[0] else None)


NFEAT = 64          # RSG_NULL_NFEAT: the surrogate's inputs (rsg_null_features)


def batch(ctx, jobs, covtype="GTp", bmin=-10.0, pmass=5e-4, fracfit=1.0, expo=True, graph=0, hists=False,
          feats=False):
    """Simulate + score + fit a batch on the device.  Jobs with group=None are nulls of one
    replicate each; jobs sharing a group id are the replicates of one null.  Returns (fits
    array (groups, 8), histograms); fits columns: mu, lambda, tau, newmass, pmass, w,
    minCOV, maxCOV.  feats=True also returns the (groups, NFEAT) surrogate features."""
    from . import Hist, Histogram
    L = _declare()
    arr = job_array(jobs)
    n = len(jobs)
    ng = int(arr["group"].max()) + 1
    fits = np.zeros(ng * 8)
    hp = (C.POINTER(Hist) * ng)() if hists else None
    fe = np.zeros((ng, NFEAT), np.float32) if feats else None
    err = _err()
    t = COV[covtype] if isinstance(covtype, str) else covtype
    _check(L.rsg_null_batch(ctx.h, _p(arr), n, t, bmin, pmass, fracfit, 1 if expo else 0,
                            C.cast(hp, C.c_void_p) if hists else None, _p(fits), _p(fe), int(graph), err),
           err, "rsg_null_batch")
    hs = [Histogram(hp[k]) if hp[k] else None for k in range(ng)] if hists else None
    if feats:
        return fits.reshape(ng, 8), hs, fe
    return fits.reshape(ng, 8), hs


def features(ctx, msa, tree):
    """The surrogate's NFEAT inputs of a resident alignment and its tree."""
    L = _declare()
    t = {k: (np.ascontiguousarray(v, dtype=np.int32) if k in ("left", "right") else
             np.ascontiguousarray(v, dtype=np.float64)) for k, v in tree.items() if k != "n"}
    f = np.zeros(NFEAT, np.float32)
    err = _err()
    _check(L.rsg_null_features(ctx.h, msa.h, _p(t["left"]), _p(t["right"]), _p(t["ld"]), _p(t["rd"]), _p(f), err),
           err, "rsg_null_features")
    return f


def evalues_tail(ctx, msa, m, lam, p0, Nc=None, matrix=True, xcap=math.inf):
    """E-values of the resident scores from a parametric tail p0 exp(-lam (x - m)), scores
    above xcap held at xcap; with matrix=False they stay on the device (for hits())."""
    L = _declare()
    n = msa.Lc
    ev = np.zeros((n, n)) if matrix else None
    err = _err()
    _check(L.rsg_cov_EvaluesTail(ctx.h, msa.h, m, lam, p0, xcap, n * (n - 1) // 2 if Nc is None else Nc, _p(ev), err),
           err, "rsg_cov_EvaluesTail")
    return ev


def hits(ctx, msa, thresh=0.05, cap=None):
    """Significant pairs of the last E-value call: arrays i, j, score, E (sorted by i, j).
    cap=None sizes the arrays to the hits (4096 first; the library reports the total when
    there are more, and the call is made again at that size); an explicit cap is a hard limit."""
    L = _declare()
    size = 4096 if cap is None else cap
    while True:
        hi = np.empty(size, np.int32); hj = np.empty(size, np.int32)
        sc = np.empty(size); ev = np.empty(size)
        n = C.c_int(0)
        err = _err()
        rc = L.rsg_cov_Hits(ctx.h, msa.h, thresh, size, C.byref(n), _p(hi), _p(hj), _p(sc), _p(ev), err)
        if rc != 0 and cap is None and n.value > size:
            size = n.value
            continue
        _check(rc, err, "rsg_cov_Hits")
        k = n.value
        return hi[:k], hj[:k], sc[:k], ev[:k]


def uneven_jobs(n=1000, seed=42, Lrange=(80, 600), Nrange=(50, 4000), blrange=(0.05, 0.3)):
    """The batched-null workload: n unstructured alignments of uneven size (L uniform over
    Lrange, N log-uniform over Nrange), each on its own Yule tree with mean branch length
    uniform over blrange and its own GTR composition, one replicate per alignment (group k)."""
    rng = np.random.default_rng(seed)
    Ls = rng.integers(Lrange[0], Lrange[1] + 1, size=n)
    Ns = np.exp(rng.uniform(np.log(Nrange[0]), np.log(Nrange[1]), size=n)).astype(int)
    bls = rng.uniform(blrange[0], blrange[1], size=n)
    jobs = []
    for k in range(n):
        pi = rng.dirichlet(np.full(4, 20.0))
        tree = yule_tree(int(Ns[k]), float(bls[k]), seed=seed * 7919 + k)
        jobs.append(Job(tree, int(Ls[k]), gtr_model(pi), seed=seed, stream=k, gapfrac=0.0))
    return jobs
