"""Train and evaluate the null-tail surrogate on the GPU.

    python -m rscape_amd.surrogate_cli train --n 50000 --out PATH.npz [--epochs E]
    python -m rscape_amd.surrogate_cli eval --model PATH.npz --n 200 [--out REPORT.json]

`train` simulates the training nulls with the batched GPU simulator, fits, trains the MLP on
the device and saves weights, scalers and training range.  `eval` draws held-out structured
alignments (planted nested pairs, as the configuration-0 generator, at random sizes), scores
each (GTp), and compares E-values of the surrogate with those of the full simulated null
(R-scape's replicate count) for pairs the full null calls significant (E < 0.05); a second
full null with other seeds gives the replicate noise of the full null itself.
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


def cmd_train(a):
    ctx = Context(0)
    t0 = time.time()
    X, Y, setups = S.generate(ctx, a.n, seed=a.seed, budget=a.budget, log=log, target_pairs=a.target_pairs)
    t1 = time.time()
    log("training pairs: %d (%.1f s to simulate, score and fit)" % (len(X), t1 - t0))
    rng = np.random.default_rng(a.seed + 1)
    idx = rng.permutation(len(X))
    nval = max(1, len(X) // 20)
    va, tr = idx[:nval], idx[nval:]
    sur = S.Surrogate(ctx, seed=a.seed)
    loss = sur.fit(X[tr], Y[tr], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed + 2)
    t2 = time.time()
    pm, pl = sur.predict(X[va])
    ok = np.isfinite(pm)
    err_u = float(np.sqrt(np.mean((pm[ok] * pl[ok] - Y[va, 0][ok]) ** 2)))
    err_l = float(np.sqrt(np.mean((np.log(pl[ok]) - Y[va, 1][ok]) ** 2)))
    const_u = float(np.sqrt(np.mean((Y[tr, 0].mean() - Y[va, 0]) ** 2)))
    const_l = float(np.sqrt(np.mean((Y[tr, 1].mean() - Y[va, 1]) ** 2)))
    sur.save(a.out)
    rep = dict(pairs=int(len(X)), simulate_s=t1 - t0, train_s=t2 - t1, final_loss=loss,
               val_rmse_lambda_m=err_u, val_rmse_lnlambda=err_l, const_rmse_lambda_m=const_u,
               const_rmse_lnlambda=const_l, epochs=a.epochs, batch=a.batch, lr=a.lr)
    log(json.dumps(rep))
    if a.report:
        json.dump(rep, open(a.report, "w"), indent=1)


def eval_one(ctx, sur, rng, Nr, Lr, K=None, seed=0, Kref=200):
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
    evs, fits = [], []
    for rep, KK in enumerate((K, K, Kref)):
        h, sf, fit = S.full_null(ctx, tree, L, pi, gap, K=KK, seed=1000 * seed + 17 + rep)
        evs.append(msa.evalues(h, sf, Nc)); fits.append(fit)
    f = NL.features(ctx, msa, tree)
    m, lam = sur.predict(f)
    res["in_range"] = bool(np.isfinite(m[0]))
    iu = np.triu_indices(L, 1)
    ef, ef2, eref = evs[0][iu], evs[1][iu], evs[2][iu]
    sig = ef < 0.05
    res["nsig"] = int(sig.sum())
    lg = lambda e: np.log10(np.maximum(e[sig], 1e-300))
    lf = lg(ef)
    res["noise"] = np.abs(lg(ef2) - lf).tolist()             # full null vs an independent full null
    res["full_vs_ref"] = np.abs(lf - lg(eref)).tolist()      # full null vs the Kref-replicate null
    res["tail_full"] = [float(fits[0][0]), float(fits[0][1])]
    res["tail_ref"] = [float(fits[2][0]), float(fits[2][1])]
    # diagnostic: the same network fed the features of one null replicate instead of the alignment's
    null_leaves = NL.simulate(ctx, NL.Job(tree, L, NL.gtr_model(pi), seed=4242 + seed, gapfrac=gap))
    mn_ = ctx.msa(null_leaves)
    fn = NL.features(ctx, mn_, tree)
    mn_.close()
    m2, lam2 = sur.predict(fn)
    if np.isfinite(m2[0]):
        es2 = NL.evalues_tail(ctx, msa, float(m2[0]), float(lam2[0]), S.P0, Nc)[iu]
        res["err_nullfeat_ref"] = np.abs(lg(es2) - lg(eref)).tolist()
        res["tail_sur_nullfeat"] = [float(m2[0]), float(lam2[0])]
        res["feat_diff"] = (f - fn).tolist()
    if res["in_range"]:
        es = NL.evalues_tail(ctx, msa, float(m[0]), float(lam[0]), S.P0, Nc)[iu]
        res["err"] = np.abs(lg(es) - lf).tolist()
        res["err_ref"] = np.abs(lg(es) - lg(eref)).tolist()
        res["agree"] = float(np.mean((es < 0.05) == sig))
        res["tail_sur"] = [float(m[0]), float(lam[0])]
        res["err_median"] = float(np.median(res["err"])) if res["err"] else None
    msa.close()
    return res


def cmd_eval(a):
    ctx = Context(0)
    sur = S.Surrogate.load(ctx, a.model)
    rng = np.random.default_rng(a.seed)
    rows = []
    for k in range(a.n):
        r = eval_one(ctx, sur, rng, (a.Nmin, a.Nmax), (a.Lmin, a.Lmax), seed=k)
        rows.append(r)
        if (k + 1) % 10 == 0:
            log("evaluated %d / %d" % (k + 1, a.n))
    def cat(key):
        return np.concatenate([np.asarray(r[key]) for r in rows if r.get(key)] or [np.zeros(0)])

    def stats(v):
        if not len(v):
            return None
        return dict(median=float(np.median(v)), p90=float(np.percentile(v, 90)), frac_within_0p3=float(np.mean(v <= 0.3)))

    out = dict(msas=len(rows), in_range=int(sum(r["in_range"] for r in rows)), sig_pairs=int(len(cat("err"))),
               surrogate_vs_full=stats(cat("err")), surrogate_vs_ref=stats(cat("err_ref")),
               surrogate_nullfeatures_vs_ref=stats(cat("err_nullfeat_ref")),
               full_vs_full=stats(cat("noise")), full_vs_ref=stats(cat("full_vs_ref")),
               call_agreement=float(np.mean([r["agree"] for r in rows if r["in_range"]])) if any(r["in_range"] for r in rows) else None)
    log(json.dumps(out))
    if a.out:
        json.dump(dict(summary=out, rows=[{k: v for k, v in r.items() if k not in ("err", "noise", "err_ref", "full_vs_ref",
                                                                                  "err_nullfeat_ref")}
                                          for r in rows]),
                  open(a.out, "w"), indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--n", type=int, default=50000)
    t.add_argument("--out", default=S.DATA)
    t.add_argument("--report", default=None)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--batch", type=int, default=512)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--budget", type=float, default=24e9)
    t.add_argument("--target-pairs", type=float, default=2e5)
    e = sub.add_parser("eval")
    e.add_argument("--model", default=S.DATA)
    e.add_argument("--n", type=int, default=100)
    e.add_argument("--seed", type=int, default=99)
    e.add_argument("--Nmin", type=int, default=20)
    e.add_argument("--Nmax", type=int, default=1000)
    e.add_argument("--Lmin", type=int, default=50)
    e.add_argument("--Lmax", type=int, default=400)
    e.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    {"train": cmd_train, "eval": cmd_eval}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
