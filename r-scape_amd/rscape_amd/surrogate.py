"""Learned null-tail surrogate.

A small MLP (67 -> 128 -> 128 -> 128 -> 2, GELU; rsg_mlp_* kernels) maps the features of an
alignment and its tree (rsg_null_features: 64) and a calibration null (3) to the exponential
tail of its phylogenetic null, written as P(score > x) = P0 * exp(-lambda (x - m)).  The
network outputs (lambda m, ln lambda): an error in lambda m moves log E by the same amount at
every score, where an error in m would be weighted by lambda, and lambda ranges over more than
an order of magnitude.  It is trained on tail fits that the batched GPU null simulator
produces (rsg_null_batch with replicate groups and features), and falls back to the full
simulated null when an alignment's inputs leave the training range.

The calibration null is a few replicates (cal_replicates: about 1e6 null pairs, at most 4)
fitted with a deeper mass (5e-3).  Past the last survival bin an E-value extrapolates the tail
by about 30 e-folds, so lambda has to be right to a few percent; two Yule trees of the same
size and mean branch length give nulls whose lambda differs by more than that (the long
branches above large clades set it), and the features of the tree alone do not pin it.  The
calibration fit sees the alignment's own tree; the network maps it, with the features, to
the tail of the full null.

Training data: random setups with N log-uniform in [20, 6000], L log-uniform in [50, 3500]
(a share `big` of them drawn from N in [1000, 6000], L in [300, 3500] instead, and a share `xl`
from N in [3000, 24000], L in [1500, 6000], the sizes of large rRNA alignments), Yule trees
with mean branch length U(0.05, 0.3), random base composition and gap fraction.  Each setup
is simulated as a group of K replicates, K = clamp(ceil(T / pairs), 1, 200) for T null pairs,
so a target fit sees T * 5e-4 pairs in its tail (fewer only for the smallest alignments); the
fit (exponential, pmass 5e-4) is moved to the reference mass P0 by
m = mu + ln(newmass / P0) / lambda.

R-scape's cov2evalue follows the fitted tail only up to the end of its 2 nb-bin survival array
(src/covariation.c:2380); every higher score gets the E-value of that last bin.  nb is set by
the largest null score, so the surrogate also predicts where the cap falls: the largest of n
null scores, in units of the tail, z = lambda (max - m), is modelled as z = a + b ln(n P0),
with (a, b) fitted on the training groups' maxima (an exponential tail would give b = 1,
a = -ln ln 2 for the median); the cap is then bmin + 2 nb w with the histogram geometry of
calculate_width_histo and cov_GrowRankList.
"""
import math
import os

import numpy as np

from . import null as NL
from .mlp import MLP

P0 = 5e-4
# Calibration null: up to CAL_MAXK simulated replicates of the alignment (about CAL_PAIRS null
# pairs), fitted with the deeper mass CAL_PMASS.  Its tail enters the network as three inputs
# (valid flag, lambda m, ln lambda at P0; zeros when the fit failed).
CAL_PMASS, CAL_PAIRS, CAL_MAXK, CAL_STREAM0 = 5e-3, 1e6, 4, 1 << 20
NCAL = 3
DIMS = (NL.NFEAT + NCAL, 128, 128, 128, 2)
CAP_DIMS = (NL.NFEAT + NCAL + 1, 64, 64, 1)     # the inputs and u = ln(null pairs x P0)
DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "surrogate_gtp.npz")


def tail_params(mu, lam, newmass):
    """(m, lambda) of P(score > x) = P0 exp(-lambda (x - m)) from an exponential tail fit."""
    return mu + math.log(newmass / P0) / lam, lam


def replicates_for(L, target_pairs=2e5, cap=200):
    pairs = L * (L - 1) // 2
    return int(min(cap, max(1, math.ceil(target_pairs / pairs))))


XL_N, XL_L = (3000, 24000), (1500, 6000)


def random_setup(rng, Nr=(20, 6000), Lr=(50, 3500), big=0.0, xl=0.0):
    u = rng.random() if (big > 0 or xl > 0) else 1.0
    if u < xl:
        Nr, Lr = XL_N, XL_L
    elif u < xl + big:
        Nr = (max(Nr[0], min(1000, Nr[1])), Nr[1]); Lr = (max(Lr[0], min(300, Lr[1])), Lr[1])
    N = int(round(math.exp(rng.uniform(math.log(Nr[0]), math.log(Nr[1])))))
    L = int(round(math.exp(rng.uniform(math.log(Lr[0]), math.log(Lr[1])))))
    pi = rng.dirichlet(np.full(4, 8.0))
    return dict(N=N, L=L, bl=float(rng.uniform(0.05, 0.3)), pi=tuple(float(p) for p in pi),
                gap=float(rng.uniform(0.0, 0.3)), tseed=int(rng.integers(1 << 31)), sseed=int(rng.integers(1 << 31)),
                tq=float(rng.random()))


def _record_bytes(L, N, K):
    return K * (48 * L * L + 3 * N * L)


def cal_replicates(L):
    pairs = L * (L - 1) // 2
    return int(min(CAL_MAXK, max(1, math.ceil(CAL_PAIRS / max(pairs, 1)))))


def cal_jobs(tree, L, model, seed, gapfrac, group=0):
    """The calibration replicates of one alignment (streams apart from the full null's)."""
    return [NL.Job(tree, L, model, seed=seed, stream=CAL_STREAM0 + r, gapfrac=gapfrac, group=group)
            for r in range(cal_replicates(L))]


def cal_inputs(fit):
    """Network inputs from a calibration fit row (mu, lambda, tau, newmass, ...)."""
    mu, lam, newmass = float(fit[0]), float(fit[1]), float(fit[3])
    if not (math.isfinite(mu) and math.isfinite(lam) and lam > 0 and newmass > 0):
        return np.zeros(NCAL, np.float32)
    m, lam = tail_params(mu, lam, newmass)
    return np.array([1.0, lam * m, math.log(lam)], np.float32)


def calibrate(ctx, tree, L, pi, gapfrac, seed=5, graph=0):
    """The calibration null of one alignment on the device: its network inputs."""
    fits, _ = NL.batch(ctx, cal_jobs(tree, L, NL.gtr_model(pi), seed, gapfrac), "GTp", pmass=CAL_PMASS,
                       expo=True, graph=graph)
    return cal_inputs(fits[0])


def inputs(ctx, msa, tree, L, pi, gapfrac, seed=5):
    """The network's inputs for a resident alignment: its features and its calibration null."""
    return np.concatenate([NL.features(ctx, msa, tree), calibrate(ctx, tree, L, pi, gapfrac, seed)])


def setup_target(st, target_pairs):
    """The null pair count a setup's group aims at: target_pairs, or for a (lo, hi) pair the
    setup's own draw, log-uniform between them (its `tq` in [0, 1))."""
    if isinstance(target_pairs, (tuple, list)):
        lo, hi = target_pairs
        return math.exp(math.log(lo) + st["tq"] * (math.log(hi) - math.log(lo)))
    return target_pairs


def generate(ctx, n, seed=1, Nr=(20, 6000), Lr=(50, 3500), budget=24e9, graph=1, log=None, covtype="GTp",
             target_pairs=2e5, big=0.0, xl=0.0):
    """n training pairs: X (n, NFEAT + NCAL) inputs (features of the group's first replicate and
    the calibration inputs of a separate calibration null), Y (n, 2) = (lambda m, ln lambda)
    of the group's fit; setups whose fit failed are dropped.  Returns X, Y, list of setups
    (each with its replicate count K, null pair count npairs, bin width w and the group's
    smallest and largest null score)."""
    rng = np.random.default_rng(seed)
    X, Y, S = [], [], []
    pending, used = [], 0.0

    def flush():
        nonlocal pending, used
        if not pending:
            return
        jobs = []
        for g, (st, tree, model) in enumerate(pending):
            K = replicates_for(st["L"], setup_target(st, target_pairs))
            jobs += [NL.Job(tree, st["L"], model, seed=st["sseed"], stream=r, gapfrac=st["gap"], group=g) for r in range(K)]
        fits, _, fe = NL.batch(ctx, jobs, covtype, pmass=5e-4, expo=True, graph=graph, feats=True)
        cj = []
        for g, (st, tree, model) in enumerate(pending):
            cj += cal_jobs(tree, st["L"], model, st["sseed"], st["gap"], group=g)
        cfits, _ = NL.batch(ctx, cj, covtype, pmass=CAL_PMASS, expo=True, graph=graph)
        for g, (st, _, _) in enumerate(pending):
            mu, lam, newmass = fits[g][0], fits[g][1], fits[g][3]
            if not (np.isfinite(mu) and np.isfinite(lam) and lam > 0 and newmass > 0):
                continue
            m, lam = tail_params(mu, lam, newmass)
            K = replicates_for(st["L"], setup_target(st, target_pairs))
            st = dict(st, K=K, npairs=K * (st["L"] * (st["L"] - 1) // 2), w=float(fits[g][5]),
                      mn=float(fits[g][6]), mx=float(fits[g][7]))
            X.append(np.concatenate([fe[g], cal_inputs(cfits[g])])); Y.append((lam * m, math.log(lam))); S.append(st)
        if log:
            log("generated %d / %d" % (len(X), n))
        pending, used = [], 0.0

    made = 0
    while made < n:
        st = random_setup(rng, Nr, Lr, big, xl)
        tree = NL.yule_tree(st["N"], st["bl"], seed=st["tseed"])
        model = NL.gtr_model(st["pi"])
        b = _record_bytes(st["L"], st["N"], max(replicates_for(st["L"], setup_target(st, target_pairs)),
                                                cal_replicates(st["L"])))
        if used + b > budget:
            flush()
        pending.append((st, tree, model)); used += b; made += 1
    flush()
    return np.array(X, np.float32), np.array(Y, np.float32), S


class Surrogate:
    """The trained network with its scalers and training range."""

    def __init__(self, ctx, net=None, seed=1):
        self.ctx = ctx
        self.net = net or MLP(ctx, DIMS, seed=seed, bf16=True)
        self.xmin = self.xmax = None
        self.zab = (-math.log(math.log(2.0)), 1.0)     # exponential tail, median maximum
        self.cap_net, self.cap_scaling, self.cap_urange = None, None, None

    @staticmethod
    def zu(setups, Y):
        """Each training group's largest null score in tail units, z = lambda (max - m), and
        u = ln(npairs P0)."""
        lam = np.exp(Y[:, 1].astype(np.float64))
        m = Y[:, 0] / lam
        z = lam * (np.array([s["mx"] for s in setups]) - m)
        u = np.log(np.array([s["npairs"] for s in setups], np.float64) * P0)
        return z, u

    def fit_cap(self, setups, Y, X=None, epochs=40, batch=256, lr=3e-4, seed=11):
        """(a, b) of z = a + b u over the training groups; with their inputs X, also a small
        network for each alignment's offset delta(inputs, u) = z - (a + b u).  One group's
        maximum carries a Gumbel spread (sd ~1.2 in z) that no input explains; under it sits an
        offset set by the tree (the split spectrum above all) and by how far past the fitted
        tail the maximum lies (u), which the network learns as a conditional mean.  Training
        groups of several sizes per alignment size (generate with a target_pairs range) give it
        u to learn from; delta() clamps u to the range it saw.  Returns (a, b)."""
        z, u = self.zu(setups, Y)
        ok = np.isfinite(z) & np.isfinite(u)
        b, a = np.polyfit(u[ok], z[ok], 1)
        self.zab = (float(a), float(b))
        if X is not None:
            d = (z - (a + b * u))[ok].astype(np.float32)[:, None]
            Xo = np.hstack([np.asarray(X, np.float32), u[:, None].astype(np.float32)])[ok]
            xm, xs = Xo.mean(0), Xo.std(0)
            xs = np.where(xs > 1e-6, xs, 1.0)
            self.cap_net = MLP(self.ctx, CAP_DIMS, seed=seed, bf16=True)
            self.cap_scaling = (xm, xs, d.mean(0), d.std(0))
            self.cap_net.set_scaling(*self.cap_scaling)
            self.cap_net.train(Xo, d, epochs=epochs, batch=batch, lr=lr, seed=seed + 1)
            self.cap_urange = (float(u[ok].min()), float(u[ok].max()))
        return self.zab

    def delta(self, feats, u):
        """The cap network's offset of z per input row at u = ln(null pairs x P0) (scalar or per
        row; clamped to the training range), 0 without a cap network."""
        feats = np.atleast_2d(np.asarray(feats, np.float32))
        if self.cap_net is None:
            return np.zeros(len(feats))
        u = np.clip(np.broadcast_to(np.asarray(u, np.float64), (len(feats),)), *self.cap_urange)
        x = np.hstack([feats, u[:, None].astype(np.float32)])
        return self.cap_net.forward(x)[:, 0].astype(np.float64)

    def tail(self, feats, K, Nc):
        """(m, lambda, xcap) of one alignment's inputs for a K-replicate null of Nc pairs; NaN,
        NaN, inf outside the training range (the caller runs the full null)."""
        mu, lam = self.predict(feats)
        if not np.isfinite(mu[0]):
            return math.nan, math.nan, math.inf
        m, lam = float(mu[0]), float(lam[0])
        delta = float(self.delta(feats, math.log(max(K * Nc * P0, 1e-300)))[0])
        return m, lam, self.cap(m, lam, K, Nc, delta=delta)

    def cap(self, m, lam, K, Nc, bmin=-10.0, w0=0.05, hpts=400, tol=1e-6, delta=0.0):
        """Score from which R-scape's E-values stop following the tail, for a K-replicate null
        of an alignment with Nc pairs: bmin + 2 nb w (src/covariation.c:2380), with w from the
        first replicate (calculate_width_histo, src/R-scape.c:1353-1360: its largest score
        against bmin) and nb = (max + 5 w - bmin) / w bins (cov_GrowRankList); the maxima
        from z = a + b u + delta."""
        a, b = self.zab
        mx1 = m + (a + b * math.log(max(Nc * P0, 1e-300)) + delta) / lam
        mx = m + (a + b * math.log(max(K * Nc * P0, 1e-300)) + delta) / lam
        w = min(w0, (mx1 - bmin) / hpts)
        if not (w >= tol and mx > bmin):
            return math.inf
        nb = int((mx + 5 * w - bmin) / w)
        return bmin + 2 * nb * w

    @staticmethod
    def base(X):
        """The calibration tail (lambda m, ln lambda) of each input row, zero where the
        calibration fit failed.  The network learns the target's difference from it: the
        calibration carries the alignment's own tree, the network the step from its mass to
        P0 and its noise, so inputs unlike the training alignments can move the prediction only
        as far as the differences it has seen."""
        X = np.atleast_2d(np.asarray(X))
        return X[:, NL.NFEAT + 1:NL.NFEAT + 3].astype(np.float64) * X[:, NL.NFEAT:NL.NFEAT + 1]

    def fit(self, X, Y, epochs=60, batch=512, lr=1e-3, seed=7):
        Yr = (np.asarray(Y, np.float64) - self.base(X)).astype(np.float32)
        xm, xs = X.mean(0), X.std(0)
        xs = np.where(xs > 1e-6, xs, 1.0)
        ym, ys = Yr.mean(0), Yr.std(0)
        self.net.set_scaling(xm, xs, ym, ys)
        self.xmin, self.xmax = X.min(0), X.max(0)
        self.net.set_range(self.xmin, self.xmax)
        self.scaling = (xm, xs, ym, ys)
        return self.net.train(X, Yr, epochs=epochs, batch=batch, lr=lr, seed=seed)

    def predict(self, feats):
        """(m, lambda) per row of inputs (inputs()); rows outside the training range get NaN."""
        feats = np.atleast_2d(np.asarray(feats, np.float32))
        out = self.net.forward(feats).astype(np.float64) + self.base(feats)
        ok = np.array([self.net.in_range(f) for f in feats])
        lam = np.exp(out[:, 1])
        m = out[:, 0] / lam
        m[~ok] = np.nan; lam[~ok] = np.nan
        return m, lam

    def save(self, path=DATA):
        W, b = self.net.get_params()
        xm, xs, ym, ys = self.scaling
        cap = {}
        if self.cap_net is not None:
            cW, cb = self.cap_net.get_params()
            cxm, cxs, cym, cys = self.cap_scaling
            cap = dict(cdims=np.array(CAP_DIMS), cxm=cxm, cxs=cxs, cym=cym, cys=cys, curange=np.array(self.cap_urange),
                       **{"cW%d" % k: w for k, w in enumerate(cW)}, **{"cb%d" % k: v for k, v in enumerate(cb)})
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, dims=np.array(DIMS), xm=xm, xs=xs, ym=ym, ys=ys, xmin=self.xmin, xmax=self.xmax,
                 zab=np.array(self.zab), **cap,
                 **{"W%d" % k: w for k, w in enumerate(W)}, **{"b%d" % k: v for k, v in enumerate(b)})

    @classmethod
    def load(cls, ctx, path=DATA):
        z = np.load(path, allow_pickle=False)
        s = cls(ctx)
        nl = len(z["dims"]) - 1
        s.net.set_params([z["W%d" % k] for k in range(nl)], [z["b%d" % k] for k in range(nl)])
        s.scaling = (z["xm"], z["xs"], z["ym"], z["ys"])
        s.net.set_scaling(*s.scaling)
        s.xmin, s.xmax = z["xmin"], z["xmax"]
        s.net.set_range(s.xmin, s.xmax)
        if "zab" in z.files:
            s.zab = (float(z["zab"][0]), float(z["zab"][1]))
        if "cdims" in z.files:
            cd = tuple(int(v) for v in z["cdims"])
            s.cap_net = MLP(ctx, cd, bf16=True)
            s.cap_net.set_params([z["cW%d" % k] for k in range(len(cd) - 1)], [z["cb%d" % k] for k in range(len(cd) - 1)])
            s.cap_scaling = (z["cxm"], z["cxs"], z["cym"], z["cys"])
            s.cap_net.set_scaling(*s.cap_scaling)
            s.cap_urange = (float(z["curange"][0]), float(z["curange"][1]))
        return s


# ------------------------------------------------------------------ significance runs
def full_null(ctx, tree, L, pi, gapfrac, K=20, seed=5, covtype="GTp", graph=0, pmass=5e-4):
    """The simulated null of one alignment: K replicates on its tree (one batched group).
    Returns (histogram, survfit, fit row)."""
    model = NL.gtr_model(pi)
    jobs = [NL.Job(tree, L, model, seed=seed, stream=r, gapfrac=gapfrac, group=0) for r in range(K)]
    fits, hs = NL.batch(ctx, jobs, covtype, pmass=pmass, expo=True, graph=graph, hists=True)
    h = hs[0]
    fit, sf = h.fit(ctx, pmass, 1.0, True)
    return h, sf, fits[0]


def full_nulls(ctx, tree, L, pi, gapfrac, seeds, K=20, covtype="GTp", pmass=5e-4):
    """len(seeds) independent simulated nulls of one alignment in one batch (one group each, the
    same jobs as full_null(seed=s) for every s): a list of (histogram, survfit, fit row)."""
    model = NL.gtr_model(pi)
    jobs = [NL.Job(tree, L, model, seed=s, stream=r, gapfrac=gapfrac, group=g)
            for g, s in enumerate(seeds) for r in range(K)]
    fits, hs = NL.batch(ctx, jobs, covtype, pmass=pmass, expo=True, hists=True)
    out = []
    for g in range(len(seeds)):
        fit, sf = hs[g].fit(ctx, pmass, 1.0, True)
        out.append((hs[g], sf, fits[g]))
    return out


def msa_null_params(ax):
    """Composition and gap fraction of an alignment (the null simulator's inputs)."""
    ax = np.asarray(ax)
    cnt = np.array([(ax == x).sum() for x in range(4)], np.float64)
    pi = cnt / cnt.sum()
    pi = np.maximum(pi, 1e-3); pi /= pi.sum()
    gap = 1.0 - cnt.sum() / ax.size
    return tuple(pi), float(gap)


def nshuffle_default(N, L):
    """R-scape's nshuffle rule (src/R-scape.c:2449-2456)."""
    ns = 20
    if N * L < 1e3:
        ns = 200
    if L < 50:
        ns = 200
    if L < 40:
        ns = 300
    if N < 4:
        ns = 300
    return ns
