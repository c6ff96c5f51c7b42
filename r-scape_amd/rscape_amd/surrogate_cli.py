"""Train and evaluate the null-tail surrogate on the GPU.

    python -m rscape_amd.surrogate_cli gen --n 100000 --out DATA.npz
    python -m rscape_amd.surrogate_cli train --data DATA.npz --out PATH.npz [--epochs E]
    python -m rscape_amd.surrogate_cli eval --model PATH.npz --n 200 [--out REPORT.json]

`gen` simulates the training nulls with the batched GPU simulator and fits their tails;
`train` (which generates when no --data is given) trains the MLP on the device, fits the
cap model on the groups' null maxima and saves weights, scalers, training range and cap
model.  `eval` draws held-out structured alignments (planted nested pairs, as the
configuration-0 generator, at random sizes), scores each (GTp), and compares the surrogate's
log10 E-values with the mean over R independent full simulated nulls (R-scape's replicate
count each) for pairs that mean calls significant (E < 0.05); the distance of one full null
from the mean of the others is the replicate noise a single R-scape run carries.  Pairs are
reported together and split into the fitted range and the capped range (past the last
survival bin, where cov2evalue holds E constant).
"""
import argparse
import json
import math
import sys
import time

import numpy as np

from . import Context
from . import null as NL
from . import surrogate as S
from . import synth


def log(msg):
    print(msg, flush=True)


SETUP_KEYS = ("N", "L", "bl", "gap", "K", "npairs", "w", "mn", "mx")


def save_data(path, X, Y, setups):
    np.savez(path, X=X, Y=Y, **{k: np.array([s[k] for s in setups], np.float64) for k in SETUP_KEYS},
             pi=np.array([s["pi"] for s in setups], np.float64))


def load_data(path):
    with np.load(path, allow_pickle=False) as z:
        arr = {k: z[k] for k in z.files}                      # each z[k] reads the archive again
    setups = [dict({k: float(arr[k][i]) for k in SETUP_KEYS}, pi=tuple(arr["pi"][i])) for i in range(len(arr["X"]))]
    return arr["X"], arr["Y"], setups


def _tp(v):
    return v[0] if len(v) == 1 else (v[0], v[1])


def cmd_gen(a):
    ctx = Context(0)
    t0 = time.time()
    X, Y, setups = S.generate(ctx, a.n, seed=a.seed, budget=a.budget, log=log, target_pairs=_tp(a.target_pairs), big=a.big,
                                  xl=a.xl)
    log("training pairs: %d (%.1f s to simulate, score and fit)" % (len(X), time.time() - t0))
    save_data(a.out, X, Y, setups)


def tail_error_at(Yt, mp, lp, depth=13.0):
    """|log10 E| error of a predicted tail at the score where the true tail has
    lambda (x - m) = depth (E ~ 1e-3 for 1e6 pairs)."""
    lt = np.exp(Yt[:, 1].astype(np.float64))
    mt = Yt[:, 0] / lt
    x = mt + depth / lt
    return np.abs(lp * (x - mp) - depth) / math.log(10.0)


def cmd_train(a):
    ctx = Context(0)
    t0 = time.time()
    if a.data:
        X, Y, setups = load_data(a.data)
    else:
        X, Y, setups = S.generate(ctx, a.n, seed=a.seed, budget=a.budget, log=log, target_pairs=_tp(a.target_pairs), big=a.big,
                                  xl=a.xl)
    t1 = time.time()
    log("training pairs: %d (%.1f s to simulate, score and fit or load)" % (len(X), t1 - t0))
    rng = np.random.default_rng(a.seed + 1)
    idx = rng.permutation(len(X))
    nval = max(1, len(X) // 20)
    va, tr = idx[:nval], idx[nval:]
    sur = S.Surrogate(ctx, seed=a.seed)
    loss = sur.fit(X[tr], Y[tr], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed + 2)
    # the cap model learns from its own data set when one is given (groups of many sizes per
    # alignment size, for the offset's dependence on u), else from the tail's
    Xc, Yc, sc, trc, vac = X, Y, setups, tr, va
    if a.cap_data:
        Xc, Yc, sc = load_data(a.cap_data)
        ic = np.random.default_rng(a.seed + 3).permutation(len(Xc))
        vac, trc = ic[:max(1, len(Xc) // 20)], ic[max(1, len(Xc) // 20):]
    zab = sur.fit_cap([sc[i] for i in trc], Yc[trc], X=Xc[trc], epochs=a.cap_epochs)
    zv, uv = S.Surrogate.zu([sc[i] for i in vac], Yc[vac])
    dv = zv - (zab[0] + zab[1] * uv)
    pv = sur.delta(Xc[vac], uv)
    okv = np.isfinite(dv)
    cap_r2 = float(1 - np.mean((dv[okv] - pv[okv]) ** 2) / np.var(dv[okv]))
    t2 = time.time()
    pm, pl = sur.predict(X[va])
    ok = np.isfinite(pm)
    err_u = float(np.sqrt(np.mean((pm[ok] * pl[ok] - Y[va, 0][ok]) ** 2)))
    err_l = float(np.sqrt(np.mean((np.log(pl[ok]) - Y[va, 1][ok]) ** 2)))
    const_u = float(np.sqrt(np.mean((Y[tr, 0].mean() - Y[va, 0]) ** 2)))
    const_l = float(np.sqrt(np.mean((Y[tr, 1].mean() - Y[va, 1]) ** 2)))
    e13 = tail_error_at(Y[va][ok], pm[ok], pl[ok])
    Nv = np.array([setups[i]["N"] for i in va])[ok]
    by_n = {}
    for lo, hi in ((20, 100), (100, 1000), (1000, 3000), (3000, 7000), (7000, 25000)):
        s = (Nv >= lo) & (Nv < hi)
        if s.any():
            by_n["N%d-%d" % (lo, hi)] = dict(n=int(s.sum()), median=float(np.median(e13[s])),
                                             p90=float(np.percentile(e13[s], 90)))
    sur.save(a.out)
    rep = dict(pairs=int(len(X)), simulate_s=t1 - t0, train_s=t2 - t1, final_loss=loss,
               val_rmse_lambda_m=err_u, val_rmse_lnlambda=err_l, const_rmse_lambda_m=const_u,
               const_rmse_lnlambda=const_l, val_log10E_err_at_depth13=by_n, cap_model_ab=list(zab),
               cap_offset_val_r2=cap_r2, cap_offset_val_sd=float(np.std(pv)),
               epochs=a.epochs, batch=a.batch, lr=a.lr)
    log(json.dumps(rep))
    if a.report:
        json.dump(rep, open(a.report, "w"), indent=1)


def eval_one(ctx, sur, rng, Nr, Lr, K=None, seed=0, R=8):
    """One held-out structured alignment.  The full null's E-value of a pair depends on the
    replicates drawn (past the last survival bin it is set by the largest null score), so the
    reference is the mean log10 E over R independent K-replicate full nulls; one full null's
    own error is its distance from the mean of the other R - 1."""
    N = int(round(math.exp(rng.uniform(math.log(Nr[0]), math.log(Nr[1])))))
    L = int(round(math.exp(rng.uniform(math.log(Lr[0]), math.log(Lr[1])))))
    d = synth.config(0, seed=int(rng.integers(1 << 31)), N=N, L=L)
    ax, tree = d["ax"], d["tree"]
    msa = ctx.msa(ax)
    msa.calculate("GTp")
    Nc = L * (L - 1) // 2
    pi, gap = S.msa_null_params(ax)
    K = K or S.nshuffle_default(N, L)
    res = dict(N=N, L=L, K=K)
    iu = np.triu_indices(L, 1)
    lg = lambda e: np.log10(np.maximum(e, 1e-300))
    lgs, caps, fits = [], [], []
    for h, sf, fit in S.full_nulls(ctx, tree, L, pi, gap, [1000 * seed + 17 + r for r in range(R)], K=K):
        lgs.append(lg(msa.evalues(h, sf, Nc)[iu])); caps.append(full_cap(h)); fits.append(fit)
    lgs = np.array(lgs)
    ref = lgs.mean(0)
    sig = ref < math.log10(0.05)
    res["nsig"] = int(sig.sum())
    cov = msa.get_cov()[0][iu]
    res["capped"] = (cov[sig] > np.median(caps)).tolist()
    loo = np.abs(lgs - (lgs.sum(0)[None] - lgs) / (R - 1))[:, sig]
    res["noise"] = loo.mean(0).tolist()                       # one full null vs the others, per pair
    res["tail_full"] = [float(np.mean([f[0] for f in fits])), float(np.mean([f[1] for f in fits]))]
    f = S.inputs(ctx, msa, tree, L, pi, gap, seed=1000 * seed + 5)
    m, lam, xcap = sur.tail(f, K, Nc)
    res["in_range"] = bool(np.isfinite(m))
    if res["in_range"]:
        es = NL.evalues_tail(ctx, msa, m, lam, S.P0, Nc, xcap=xcap)[iu]
        res["err"] = np.abs(lg(es[sig]) - ref[sig]).tolist()
        res["agree"] = float(np.mean((es < 0.05) == sig))
        res["tail_sur"] = [m, lam]
        res["err_median"] = float(np.median(res["err"])) if res["err"] else None
    msa.close()
    return res


def full_cap(h):
    """Score above which a full null's E-values are constant: cov2evalue reads
    survfit[min(icov + 1, 2 nb - 1)] (src/covariation.c:2380-2381)."""
    g = h.h
    return g.bmin + (2 * g.nb - 2) * g.w


def cmd_eval(a):
    ctx = Context(0)
    sur = S.Surrogate.load(ctx, a.model)
    rng = np.random.default_rng(a.seed)
    rows = []
    for k in range(a.n):
        r = eval_one(ctx, sur, rng, (a.Nmin, a.Nmax), (a.Lmin, a.Lmax), seed=k, R=a.nulls)
        rows.append(r)
        if (k + 1) % 10 == 0:
            log("evaluated %d / %d" % (k + 1, a.n))
    def cat(key, part=None):
        out = []
        for r in rows:
            if not r.get(key):
                continue
            v = np.asarray(r[key])
            if part is not None:
                c = np.asarray(r["capped"], bool)
                v = v[c] if part == "capped" else v[~c]
            out.append(v)
        return np.concatenate(out) if out else np.zeros(0)

    def stats(v):
        if not len(v):
            return None
        return dict(median=float(np.median(v)), p90=float(np.percentile(v, 90)), frac_within_0p3=float(np.mean(v <= 0.3)))

    out = dict(model=a.model, msas=len(rows), in_range=int(sum(r["in_range"] for r in rows)), sig_pairs=int(len(cat("err"))),
               reference="mean log10 E over %d independent full nulls of R-scape's replicate count" % a.nulls,
               surrogate_vs_full=stats(cat("err")), one_full_null=stats(cat("noise")),
               fit_range=dict(pairs=int(len(cat("err", "fit"))), surrogate_vs_full=stats(cat("err", "fit")),
                              one_full_null=stats(cat("noise", "fit"))),
               capped=dict(pairs=int(len(cat("err", "capped"))), surrogate_vs_full=stats(cat("err", "capped")),
                           one_full_null=stats(cat("noise", "capped"))),
               call_agreement=float(np.mean([r["agree"] for r in rows if r["in_range"]])) if any(r["in_range"] for r in rows) else None)
    log(json.dumps(out))
    if a.out:
        json.dump(dict(summary=out, rows=[{k: v for k, v in r.items() if k not in ("err", "noise", "capped")}
                                          for r in rows]),
                  open(a.out, "w"), indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gen")
    g.add_argument("--n", type=int, default=100000)
    g.add_argument("--out", required=True)
    g.add_argument("--seed", type=int, default=1)
    g.add_argument("--budget", type=float, default=24e9)
    g.add_argument("--target-pairs", type=float, nargs="+", default=[2e6],
                   help="null pairs per training group, or a range LO HI drawn log-uniformly per setup")
    g.add_argument("--big", type=float, default=0.5)
    g.add_argument("--xl", type=float, default=0.03)
    t = sub.add_parser("train")
    t.add_argument("--data", default=None)
    t.add_argument("--n", type=int, default=100000)
    t.add_argument("--big", type=float, default=0.5)
    t.add_argument("--xl", type=float, default=0.03)
    t.add_argument("--out", default=S.DATA)
    t.add_argument("--report", default=None)
    t.add_argument("--epochs", type=int, default=120)
    t.add_argument("--cap-epochs", type=int, default=40)
    t.add_argument("--cap-data", default=None, help="training data of the cap model (default: --data)")
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=3e-4)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--budget", type=float, default=24e9)
    t.add_argument("--target-pairs", type=float, nargs="+", default=[2e6])
    e = sub.add_parser("eval")
    e.add_argument("--model", default=S.DATA)
    e.add_argument("--n", type=int, default=100)
    e.add_argument("--seed", type=int, default=99)
    e.add_argument("--Nmin", type=int, default=20)
    e.add_argument("--Nmax", type=int, default=1000)
    e.add_argument("--Lmin", type=int, default=50)
    e.add_argument("--Lmax", type=int, default=400)
    e.add_argument("--nulls", type=int, default=8)
    e.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    {"gen": cmd_gen, "train": cmd_train, "eval": cmd_eval}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
