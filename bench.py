#!/usr/bin/env python3
"""Benchmark of the pairwise-covariation hot path on one MI355X.

One step = one complete scoring of the resident alignment as R-scape's cov_Calculate
does it for GTp (src/covariation.c:63-306): pair counts (int8 one-hot MFMA Gram),
probabilities and marginals (FP64), the GT statistic (FP32 logarithms on FP64 probabilities,
within 1e-5 of FP64; --stat fp64 for the FP64 statistic), the APC correction and the score
extremes.  Workload: synthetic rRNA-like alignment N=5000 sequences x L=1500
columns (400 planted nested pairs, 10% gaps in runs).  Metric: column-pair x sequence
updates per second, updates = N * L(L-1)/2 per step.

Prints ONE JSON line.  Extra fields report the weighted variants (int8 weight limbs and
split-BF16), the FP32 and FP64 statistics, a per-kernel time breakdown, the roofline of the
dominant kernel (the int8 Gram), the CPU oracle on the same alignment, complete significance
runs (per-replicate null, batched null, surrogate; their E-value accuracy), the batched null of
1000 uneven alignments, and the N=20000 x L=5000 weighted significance run (--no-evalue,
--no-batch, --no-c4 skip the last three).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "r-scape_amd"))

import numpy as np  # noqa: E402

import rscape_amd as R  # noqa: E402
from rscape_amd import synth  # noqa: E402

PEAK_INT8_TOPS = 5000.0      # MI355X dense int8 MFMA, TOP/s
PEAK_BF16_TFLOPS = 2500.0    # MI355X dense bf16 MFMA, TFLOP/s
PEAK_HBM_GBS = 8000.0
METRIC = "column-pair\u00b7seq updates/s (GT+APC); E-value runs/s, surrogate vs full null"


def kernel_avgs(timings):
    acc = {}
    for name, ms in timings:
        acc.setdefault(name, []).append(ms)
    return {k: (float(np.mean(v)), len(v)) for k, v in acc.items()}


def run_steps(ctx, m, k, breakdown=5, **kw):
    """k timed scorings with HIP events around the Gram launches only (its live duration for
    the roofline), then `breakdown` more with events around every stage (kernels_ms)."""
    ctx.set_timing("gram")
    t0 = time.perf_counter()
    for _ in range(k):
        m.calculate(**kw)
    t1 = time.perf_counter()
    gram = kernel_avgs(ctx.timings())
    ctx.set_timing(True)
    for _ in range(breakdown):
        m.calculate(**kw)
    allk = kernel_avgs(ctx.timings())
    ctx.set_timing(False)
    allk.update(gram)
    return t1 - t0, allk


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return min(16, n)                # one GPU's share of the host (16 cores)


def cpu_baseline(ax, Lwin, threads=1):
    """Oracle (FP64 restatement of the reference loops: counts, probabilities, marginals, GT,
    APC) on the first Lwin columns, with `threads` threads over the column-pair rows (the
    reference's own parallelism is independent runs over replicates or alignments; the rows of
    one scoring are as independent)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    try:
        import oracle
        oracle.lib()
    except Exception as e:  # the oracle library is optional where the benchmark runs
        return {"value": None, "error": str(e)}
    sub = np.ascontiguousarray(ax[:, :Lwin])
    N = sub.shape[0]
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    pr = oracle.probs(sub)
    cv, _ = oracle.cov(pr, oracle.GT, oracle.CSELECT)
    oracle.corrected(cv, Lwin, oracle.APC)
    dt = time.perf_counter() - t0
    oracle.set_threads(1)
    upd = N * Lwin * (Lwin - 1) / 2
    return {"value": upd / dt, "unit": "column-pair*seq updates/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": "GT+APC (counts, probabilities, marginals, GT, APC) on columns 0..%d of the same "
                      "N=%d alignment, %.1f s, %d thread(s)" % (Lwin - 1, N, dt, threads)}


def uneven_null_batch(ctx, reps=3):
    """The batched null of 1000 unstructured alignments of uneven size (L uniform 80-600, N
    log-uniform 50-4000, own Yule trees, mean branch length 0.05-0.3): simulate, score GTp,
    histogram and exponential tail fit per alignment, all in one call; eager and as one captured
    graph.  The graph is captured by the first call (timed apart as `first_call_s`); each timed
    call draws new seeds for the same sizes and trees and replays it."""
    from rscape_amd import null as NL
    jobs = NL.uneven_jobs(1000, seed=42)
    upd = float(sum(j.tree["n"] * j.L * (j.L - 1) / 2 for j in jobs))
    out = {"workload": "1000 alignments, L ~ U(80,600), N ~ logU(50,4000), GTp, one group each",
           "updates": upd}
    for name, graph, persistent in (("eager_per_tile", 0, False), ("eager", 0, True), ("graph", 1, True)):
        ctx.set_batch_persistent(persistent)
        t0 = time.perf_counter()
        NL.batch(ctx, jobs, "GTp", pmass=5e-4, graph=graph)
        first = time.perf_counter() - t0
        ts = []
        for r in range(reps):
            for j in jobs:
                j.seed += 1
            t0 = time.perf_counter()
            NL.batch(ctx, jobs, "GTp", pmass=5e-4, graph=graph)
            ts.append(time.perf_counter() - t0)
        t = float(np.median(ts))
        out[name] = {"s_per_batch": t, "alignments_per_s": 1000.0 / t, "updates_per_s": upd / t,
                     "first_call_s": first}
    ctx.set_batch_persistent(True)
    out["graph_reuses"] = ctx.graph_reuses()
    # where one eager call's time goes: the job records built in Python, the library's host
    # planning, the uploads, and the device run (launch to completion)
    t0 = time.perf_counter()
    NL.job_array(jobs)
    phases = {"python_job_records": (time.perf_counter() - t0) * 1e3}
    ctx.set_timing(True)
    NL.batch(ctx, jobs, "GTp", pmass=5e-4, graph=0)
    for name, ms in ctx.timings():
        phases[name] = phases.get(name, 0.0) + ms
    ctx.set_timing(False)
    out["phases_ms"] = phases
    return out


def significance_accuracy(ctx, m, tree, L, pi, gap, K, R):
    """The reference for E-value errors of a scored alignment m (GTp): the full null's E-value of
    a pair is itself random (it depends on the replicates drawn), most of all past the last
    survival bin, where it is set by the largest null score (src/covariation.c:2380).  The
    reference log10 E is therefore the mean over R independent K-replicate full nulls; one full
    null's own error is its distance from the mean of the other R - 1 (the replicate noise any
    single run of R-scape carries).  Returns (fields, fn) with fn(E matrix) -> the fields of
    that matrix's errors on the pairs the reference calls (E < 0.05), split into the fitted
    range and the capped range (past the last survival bin)."""
    from rscape_amd import surrogate as S
    Nc = L * (L - 1) // 2
    iu = np.triu_indices(L, 1)
    lg = lambda e: np.log10(np.maximum(e, 1e-300)).astype(np.float32)
    cov = m.get_cov()[0][iu]
    lgs, capped_from = np.empty((R, len(iu[0])), np.float32), []
    for r in range(R):
        h, sf, _ = S.full_null(ctx, tree, L, pi, gap, K=K, seed=1000 + r)
        lgs[r] = lg(m.evalues(h, sf, Nc)[iu])
        capped_from.append(h.h.bmin + (2 * h.h.nb - 2) * h.h.w)
    ref = lgs.mean(0)
    sig = ref < math.log10(0.05)
    capped = sig & (cov > np.median(capped_from))
    fitr = sig & ~capped

    def errs(v, sel):
        v = v[sel]
        return dict(pairs=int(sel.sum()), median=float(np.median(v)) if len(v) else None,
                    frac_within_0p3=float(np.mean(v <= 0.3)) if len(v) else None)
    loo = np.abs(lgs[:, sig] - (lgs[:, sig].sum(0)[None] - lgs[:, sig]) / (R - 1)).reshape(-1)
    til = lambda s: np.tile(s[sig], R)
    out = {"accuracy_reference": "mean log10 E over %d independent %d-replicate full nulls" % (R, K),
           "sig_pairs": int(sig.sum()),
           "one_full_null_log10E": {"all": errs(loo, np.ones(len(loo), bool)), "fit_range": errs(loo, til(fitr)),
                                    "capped": errs(loo, til(capped))}}

    def of(ev):
        d = np.abs(lg(ev[iu]) - ref)
        return {"surrogate_log10E": {"all": errs(d, sig), "fit_range": errs(d, fitr), "capped": errs(d, capped)},
                "call_agreement": float(np.mean((ev[iu] < 0.05) == sig))}
    return out, of


def fitted_range_variant(ctx, N=200, K=20, R=64):
    """Surrogate accuracy on the headline's shape with few sequences (N=200 x L=1500, 400
    planted pairs): the planted pairs score inside the fitted tail instead of past the last
    survival bin, so this tests the tail the network predicts rather than the cap."""
    from rscape_amd import null as NL
    from rscape_amd import surrogate as S
    d = synth.config(1, seed=42, N=N)
    ax, tree = d["ax"], d["tree"]
    L = ax.shape[1]
    Nc = L * (L - 1) // 2
    pi, gap = S.msa_null_params(ax)
    m = ctx.msa(ax)
    m.calculate("GTp")
    out, of = significance_accuracy(ctx, m, tree, L, pi, gap, K, R)
    out["workload"] = "N=%d x L=%d, GTp, %d planted pairs, %d-replicate nulls" % (N, L, int((d["ct"] >= 0).sum()) // 2, K)
    if os.path.exists(S.DATA):
        sur = S.Surrogate.load(ctx)
        mu, lam, xcap = sur.tail(S.inputs(ctx, m, tree, L, pi, gap, seed=7), K, Nc)
        if np.isfinite(mu):
            out.update(of(NL.evalues_tail(ctx, m, mu, lam, S.P0, Nc, matrix=True, xcap=xcap)))
        else:
            out["surrogate_fallback"] = True
    m.close()
    return out


def evalue_runs(ctx, d, K=20, reps=3, acc_nulls=64):
    """Complete significance runs of the alignment (GTp scoring + null + exponential tail fit +
    E-values + the list of pairs with E < 0.05), three ways: the plain per-replicate GPU path
    (simulate one replicate, score it, bin it, merge on the host: R-scape's loop with the GPU
    doing each step), the batched null (all K replicates in one call), and the learned
    surrogate (features + MLP + parametric tail; full null when the features leave the
    training range).  E-value errors are measured outside the timed runs."""
    from rscape_amd import null as NL
    from rscape_amd import surrogate as S
    ax, tree = d["ax"], d["tree"]
    N, L = ax.shape
    Nc = L * (L - 1) // 2
    pi, gap = S.msa_null_params(ax)
    model = NL.gtr_model(pi)
    m = ctx.msa(ax)
    sur = S.Surrogate.load(ctx) if os.path.exists(S.DATA) else None
    fallback = [False]

    def plain(matrix=False):
        m.calculate("GTp")
        cum, w = None, None
        for r in range(K):
            leaves = NL.simulate(ctx, NL.Job(tree, L, model, seed=7, stream=r, gapfrac=gap))
            mr = ctx.msa(leaves)
            a, b = mr.calculate("GTp")
            if w is None:
                w = R.width(b, a)
            cum = R.Histogram.null_add(cum, mr.rank_hist(-10.0, w))
            mr.close()
        _, sf = cum.fit(ctx, 5e-4, 1.0, True)
        ev = m.evalues(cum, sf, Nc, matrix=matrix)
        return ev if matrix else NL.hits(ctx, m, 0.05)

    def batched(matrix=False, graph=0):
        m.calculate("GTp")
        h, sf, _ = S.full_null(ctx, tree, L, pi, gap, K=K, seed=7, graph=graph)
        ev = m.evalues(h, sf, Nc, matrix=matrix)
        return ev if matrix else NL.hits(ctx, m, 0.05)

    seeds = [7 + r for r in range(K)]

    def plain_fitch(matrix=False):
        # R-scape's own null (null_rscape): Fitch ancestors + column shuffle per replicate
        m.calculate("GTp")
        cum, w = None, None
        for sd in seeds:
            mr = ctx.msa(NL.fitch_null(ctx, ax, tree, seed=sd))
            a, b = mr.calculate("GTp")
            if w is None:
                w = R.width(b, a)
            cum = R.Histogram.null_add(cum, mr.rank_hist(-10.0, w))
            mr.close()
        _, sf = cum.fit(ctx, 5e-4, 1.0, True)
        ev = m.evalues(cum, sf, Nc, matrix=matrix)
        return ev if matrix else NL.hits(ctx, m, 0.05)

    def batched_fitch(matrix=False, graph=0):
        m.calculate("GTp")
        _, h = NL.batch_fitch(ctx, ax, tree, seeds, "GTp", pmass=5e-4, graph=graph)
        _, sf = h.fit(ctx, 5e-4, 1.0, True)
        ev = m.evalues(h, sf, Nc, matrix=matrix)
        return ev if matrix else NL.hits(ctx, m, 0.05)

    def surrogate(matrix=False):
        m.calculate("GTp")
        mu, lam, xcap = sur.tail(S.inputs(ctx, m, tree, L, pi, gap, seed=7), K, Nc)
        if not np.isfinite(mu):
            fallback[0] = True
            return batched(matrix)
        ev = NL.evalues_tail(ctx, m, mu, lam, S.P0, Nc, matrix=matrix, xcap=xcap)
        return ev if matrix else NL.hits(ctx, m, 0.05)

    def timed(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    out = {"unit": "runs/s", "replicates": K,
           "workload": "N=%d x L=%d, GTp, exponential tail, pmass 5e-4, hit list E < 0.05" % (N, L)}
    tp = timed(plain)
    tb = timed(batched)
    tg = timed(lambda: batched(graph=1))
    out["plain_per_replicate"] = 1.0 / tp
    out["batched_null"] = 1.0 / tb
    out["batched_null_graph"] = 1.0 / tg
    out["speedup_batched_vs_plain"] = tp / tb
    tpf = timed(plain_fitch)
    tbf = timed(lambda: batched_fitch(graph=1))
    out["rscape_null_plain_per_replicate"] = 1.0 / tpf
    out["rscape_null_batched_graph"] = 1.0 / tbf
    out["speedup_rscape_null_batched_vs_plain"] = tpf / tbf
    ef, eg = plain_fitch(matrix=True), batched_fitch(matrix=True, graph=1)
    ok = np.isfinite(ef) & np.isfinite(eg)
    out["rscape_null_plain_vs_batched_max_abs_log10E"] = float(np.max(np.abs(
        np.log10(np.maximum(ef[ok], 1e-300)) - np.log10(np.maximum(eg[ok], 1e-300)))))
    iu = np.triu_indices(L, 1)
    eb = batched(matrix=True)[iu]
    ep = plain(matrix=True)[iu]
    sig = eb < 0.05
    lg = lambda e: np.log10(np.maximum(e, 1e-300))
    out["plain_vs_batched_max_abs_log10E"] = float(np.max(np.abs(lg(ep[sig]) - lg(eb[sig])))) if sig.any() else 0.0
    m.calculate("GTp")
    acc, ref_sig = significance_accuracy(ctx, m, tree, L, pi, gap, K, acc_nulls)
    out.update(acc)
    if sur is not None:
        ts = timed(surrogate)
        out["surrogate"] = 1.0 / ts
        out["surrogate_fallback"] = fallback[0]
        out["speedup_surrogate_vs_plain"] = tp / ts
        out["speedup_surrogate_vs_rscape_null_plain"] = tpf / ts
        out["speedup_surrogate_vs_rscape_null_batched"] = tbf / ts
        out.update(ref_sig(surrogate(matrix=True)))
    m.close()
    return out


def c4_significance(ctx, K=100, R_nulls=32, reps=3):
    """Config 4 end to end: the N=20000 x L=5000 alignment (1200 nested + 40 pseudoknotted pairs,
    15% gaps), position-based weights, weighted GT+APC scoring (int8-limb counts, FP32
    statistics), then its significance three ways -- the plain per-replicate GPU path (R-scape's
    loop), one batch of K replicates, and the surrogate -- each a complete run: scoring, null,
    fit, E-values, hit list (E < 0.05).  The nulls score unit-weight GTp on simulated alignments
    of the same size down the alignment's tree.  E-value accuracy is measured against the mean
    log10 E over R independent K-replicate full nulls (a full null's E-value is itself random),
    with one full null's distance from the mean of the others beside it."""
    from rscape_amd import null as NL
    from rscape_amd import surrogate as S
    tstart = t0 = time.perf_counter()
    d = synth.config(4, seed=42)
    ax, tree = d["ax"], d["tree"]
    N, L = ax.shape
    Nc = L * (L - 1) // 2
    out = {"workload": "N=%d x L=%d, PB weights, GTp (W8 counts, FP32 statistic), %d-replicate null" % (N, L, K),
           "generate_s": time.perf_counter() - t0}
    say = lambda msg: print("[c4] %s (%.1f s)" % (msg, time.perf_counter() - tstart), file=sys.stderr, flush=True)
    say("alignment generated")
    w = ctx.weights(ax, "PB")
    m = ctx.msa(ax, w)
    kw = dict(covtype="GTp", counts=R.COUNTS_W8, prec=R.STAT_FP32)
    m.calculate(**kw)
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        m.calculate(**kw)
        ts.append(time.perf_counter() - t0)
    say("weighted scoring timed")
    out["weighted_scoring_ms"] = float(np.median(ts)) * 1e3
    out["weighted_scoring_updates_per_s"] = N * L * (L - 1) / 2 / float(np.median(ts))
    # the int8-limb Gram alone (HIP events around its launches): 3 limbs x 16 int8 MACs per update
    ctx.set_timing("gram")
    for _ in range(3):
        m.calculate(**kw)
    tg = [ms for name, ms in ctx.timings() if name == "gram_w8"]
    ctx.set_timing(False)
    if tg:
        gms = float(np.mean(tg))
        out["weighted_gram_w8_ms"] = gms
        out["weighted_gram_w8_frac_int8_peak"] = 3 * 32 * N * L * (L - 1) / 2 / (gms * 1e-3) / (PEAK_INT8_TOPS * 1e12)
    pi, gap = S.msa_null_params(ax)
    model = NL.gtr_model(pi)
    sur = S.Surrogate.load(ctx) if os.path.exists(S.DATA) else None

    def plain():
        m.calculate(**kw)
        cum, wd = None, None
        for r in range(K):
            mr = ctx.msa(NL.simulate(ctx, NL.Job(tree, L, model, seed=7, stream=r, gapfrac=gap)))
            a, b = mr.calculate("GTp")
            if wd is None:
                wd = R.width(b, a)
            cum = R.Histogram.null_add(cum, mr.rank_hist(-10.0, wd))
            mr.close()
        _, sf = cum.fit(ctx, 5e-4, 1.0, True)
        return m.evalues(cum, sf, Nc, matrix=True)

    def batched():
        m.calculate(**kw)
        h, sf, _ = S.full_null(ctx, tree, L, pi, gap, K=K, seed=7)
        return m.evalues(h, sf, Nc, matrix=True)

    def surrogate():
        m.calculate(**kw)
        mu, lam, xcap = sur.tail(S.inputs(ctx, m, tree, L, pi, gap, seed=7), K, Nc)
        if not np.isfinite(mu):
            return batched(), True
        return NL.evalues_tail(ctx, m, mu, lam, S.P0, Nc, matrix=True, xcap=xcap), False

    def timed(fn, n=reps):
        r = fn()                                 # the first call allocates the buffers
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t0)
        return r, float(np.median(ts))

    iu = np.triu_indices(L, 1)
    lg = lambda e: np.log10(np.maximum(e, 1e-300)).astype(np.float32)
    eb, tb = timed(batched)
    say("batched null timed")
    ep, tp = timed(plain, 1)
    say("plain per-replicate null timed")
    out["plain_per_replicate_s"] = tp
    out["batched_null_s"] = tb
    out["speedup_batched_vs_plain"] = tp / tb
    eb, ep = lg(eb[iu]), lg(ep[iu])
    out["plain_vs_batched_max_abs_log10E"] = float(np.max(np.abs(ep - eb)))
    del ep
    if sur is not None:
        (es, fb), tsur = timed(surrogate)
        say("surrogate timed")
        out["surrogate_s"] = tsur
        out["surrogate_fallback"] = fb
        out["speedup_surrogate_vs_plain"] = tp / tsur
        out["speedup_surrogate_vs_batched"] = tb / tsur
        es = lg(es[iu])
    # accuracy reference: R independent full nulls (other seeds than the timed ones)
    lgs = np.empty((R_nulls, len(iu[0])), np.float32)
    m.calculate(**kw)
    for r in range(R_nulls):
        h, sf, _ = S.full_null(ctx, tree, L, pi, gap, K=K, seed=1000 + r)
        lgs[r] = lg(m.evalues(h, sf, Nc, matrix=True)[iu])
    say("%d reference nulls" % R_nulls)
    ref = lgs.mean(0)
    sig = ref < math.log10(0.05)
    out["sig_pairs"] = int(sig.sum())
    planted = np.zeros((L, L), bool)            # ct holds the nested and the pseudoknotted pairs
    for i, j in enumerate(d["ct"]):
        if j > i:
            planted[i, j] = True
    out["planted_pairs"] = int(planted.sum())
    out["planted_called"] = int((planted[iu] & sig).sum())

    def err(v):
        dd = np.abs(v[sig] - ref[sig]).astype(np.float64)
        return {"median": float(np.median(dd)), "p90": float(np.percentile(dd, 90)),
                "frac_within_0p3": float(np.mean(dd <= 0.3))} if sig.any() else None
    loo = np.abs(lgs[:, sig] - (lgs[:, sig].sum(0)[None] - lgs[:, sig]) / (R_nulls - 1))
    out["accuracy_reference"] = "mean log10 E over %d independent %d-replicate full nulls" % (R_nulls, K)
    out["one_full_null_log10E"] = {"median": float(np.median(loo)), "frac_within_0p3": float(np.mean(loo <= 0.3))}
    out["batched_null_log10E"] = err(eb)
    if sur is not None:
        out["surrogate_log10E"] = err(es)
        out["call_agreement"] = float(np.mean((es < math.log10(0.05)) == sig))
    m.close()
    return out


def fp32_errors(ctx, ax):
    """Max relative error of each FP32 statistic against the FP64 path on the same counts
    (floored at 1e-3 of the largest |score|, as the parity tests do)."""
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    L = ax.shape[1]
    iu = np.triu_indices(L, 1)
    out = {}
    for name in ("GT", "MI", "MIr", "MIg", "CHI"):
        m.cov(name, R.C16, prec=R.STAT_FP64)
        ref = m.get_cov()[0][iu]
        m.cov(name, R.C16, prec=R.STAT_FP32)
        got = m.get_cov()[0][iu]
        floor = 1e-3 * np.max(np.abs(ref))
        out[name] = float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor)))
    m.close()
    return out


def load_traffic(name, key="gram_traffic_bytes_per_scoring:"):
    """HBM bytes of the Gram (both of its launches) from a committed rocprofv3 counter summary
    (or None)."""
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        for line in open(p):
            if line.startswith(key):
                return float(line.split(":")[1])
    return None


CALIB_BYTES = 256 << 20


def calibration_kernel(ctx):
    """The library's plain copy kernel (k_copy_probe: reads and writes CALIB_BYTES per launch),
    run outside the timed region so rocprofv3 counter runs can convert FETCH_SIZE /
    WRITE_SIZE to bytes."""
    import ctypes
    err = ctypes.create_string_buffer(128)
    L = R.lib()
    L.rsg_ctx_copy_probe.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_char_p]
    if L.rsg_ctx_copy_probe(ctx.h, CALIB_BYTES, err) != 0:
        raise RuntimeError(err.value.decode())


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # enough untimed steps for the GPU clocks to leave their idle state after the host-side data
    # generation, and enough timed ones (~60 ms) for a steady rate
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--N", type=int, default=5000)
    ap.add_argument("--L", type=int, default=1500)
    ap.add_argument("--cpu-window", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--calib", action="store_true", help="also run the counter calibration kernel")
    ap.add_argument("--no-evalue", action="store_true", help="skip the significance-run timing")
    ap.add_argument("--evalue-reps", type=int, default=3)
    ap.add_argument("--stat", choices=("fp32", "fp64"), default="fp32",
                    help="arithmetic of the per-pair statistics in the timed step")
    ap.add_argument("--no-batch", action="store_true", help="skip the 1000-alignment null batch")
    ap.add_argument("--c4", action="store_true",
                    help="only the N=20000 x L=5000 significance run (config 4: weighted scoring, 100-replicate null, surrogate)")
    ap.add_argument("--no-c4", action="store_true", help="skip the N=20000 x L=5000 significance run")
    ap.add_argument("--null-batch-only", action="store_true",
                    help="time only the 1000-alignment null batch and print its JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(a sharded design is described with the library). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.null_batch_only:
        ctx = R.Context(0)
        print(json.dumps({"uneven_null_batch": uneven_null_batch(ctx)}), flush=True)
        ctx.close()
        return
    if args.c4:
        ctx = R.Context(0)
        print(json.dumps({"config4_significance": c4_significance(ctx)}), flush=True)
        ctx.close()
        return

    d = synth.config(1, seed=42, N=args.N, L=args.L)
    ax = d["ax"]
    N, L = ax.shape
    upd = N * L * (L - 1) / 2

    ctx = R.Context(0)
    m = ctx.msa(ax)
    prec = R.STAT_FP32 if args.stat == "fp32" else R.STAT_FP64
    kw = dict(covtype="GTp", counts=R.COUNTS_INT8, prec=prec)
    for _ in range(args.warmup):
        m.calculate(**kw)
    dt, kav = run_steps(ctx, m, args.steps, **kw)
    value = upd * args.steps / dt
    # the int8 Gram is k_gram_i8 plus, when the last partial wave of tiles is split along K,
    # k_gram_fix: one pair of HIP events on the library stream brackets both launches
    g_ms = kav.get("gram_i8", (float("nan"), 0))[0]
    gram_ops = 2.0 * 16.0 * upd                      # 16 int8 MACs per update (4x4 canonical table)
    achieved = gram_ops / (g_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "N=%d x L=%d synthetic rRNA-like MSA, GTp (unit weights, int8 counts, %s "
                               "statistics), one full scoring per step" % (N, L, args.stat.upper()),
                   "N": N, "L": L, "covtype": "GTp", "counts": "int8 one-hot Gram", "statistic": args.stat.upper()},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_INT8_TOPS, "unit": "TOP/s",
                     "frac": achieved / PEAK_INT8_TOPS,
                     "traffic": load_traffic("gram_i8_n5000_l1500_traffic.txt"),
                     "kernel": "k_gram_i8 + k_gram_fix", "kernel_ms": g_ms,
                     "algorithmic_ops_per_launch": gram_ops},
        "kernels_ms": {k: v[0] for k, v in kav.items()},
    }
    if not args.no_extra:
        extra = {}
        m.close()
        rng = np.random.default_rng(1)
        w = rng.uniform(0.2, 2.0, size=N)
        mw = ctx.msa(ax, w)
        for sname, sprec in (("fp32", R.STAT_FP32), ("fp64", R.STAT_FP64)):
            kw2 = dict(covtype="GTp", counts=R.COUNTS_BF16, prec=sprec)
            mw.calculate(**kw2)
            dtw, kw_av = run_steps(ctx, mw, max(3, args.steps // 4), **kw2)
            b_ms = kw_av.get("gram_bf16", (float("nan"), 0))[0]
            extra["weighted_bf16_hilo_" + sname + "_statistic"] = {
                "value": upd * max(3, args.steps // 4) / dtw, "unit": "updates/s",
                "gram_bf16_ms": b_ms, "gram_bf16_tflops": 2 * 2 * 16 * upd / (b_ms * 1e-3) / 1e12,
                "peak_tflops": PEAK_BF16_TFLOPS, "stat_ms": kw_av.get("stat", (float("nan"), 0))[0]}
        # exact integer weighted counts: 21-bit weights in three 7-bit limbs through the int8 Gram
        kw2 = dict(covtype="GTp", counts=R.COUNTS_W8, prec=R.STAT_FP32)
        mw.calculate(**kw2)
        dtw, kw_av = run_steps(ctx, mw, max(3, args.steps // 4), **kw2)
        w_ms = kw_av.get("gram_w8", (float("nan"), 0))[0]
        extra["weighted_int8_limbs_fp32_statistic"] = {
            "value": upd * max(3, args.steps // 4) / dtw, "unit": "updates/s",
            "gram_w8_ms": w_ms, "gram_w8_tops": 3 * 32 * upd / (w_ms * 1e-3) / 1e12,
            "peak_tops": PEAK_INT8_TOPS, "stat_ms": kw_av.get("stat", (float("nan"), 0))[0]}
        mw.close()
        m = ctx.msa(ax)
        other = R.STAT_FP64 if prec == R.STAT_FP32 else R.STAT_FP32
        kw3 = dict(covtype="GTp", counts=R.COUNTS_INT8, prec=other)
        m.calculate(**kw3)
        dt3, kav3 = run_steps(ctx, m, max(3, args.steps // 4), **kw3)
        extra["fp64_statistic" if other == R.STAT_FP64 else "fp32_statistic"] = {
            "value": upd * max(3, args.steps // 4) / dt3, "unit": "updates/s",
            "stat_ms": kav3.get("stat", (float("nan"), 0))[0]}
        extra["fp32_max_rel_err_vs_fp64"] = fp32_errors(ctx, ax)
        # PCIe-inclusive: the alignment uploaded from host memory before every scoring
        m.calculate(**kw)
        k = max(3, args.steps // 4)
        t0 = time.perf_counter()
        for _ in range(k):
            m.update(ax)
            m.calculate(**kw)
        extra["pcie_inclusive"] = {"value": upd * k / (time.perf_counter() - t0), "unit": "updates/s",
                                   "note": "N x L codes copied host->device and re-packed before each step"}
        out["extra"] = extra
    if not args.no_evalue:
        out["evalue_runs"] = evalue_runs(ctx, d, reps=args.evalue_reps)
        out["evalue_runs"]["fitted_range_variant"] = fitted_range_variant(ctx)
    if not args.no_batch:
        out["uneven_null_batch"] = uneven_null_batch(ctx)
    if not args.no_c4:
        out["config4_significance"] = c4_significance(ctx)
    if args.calib:
        calibration_kernel(ctx)
    if not args.no_cpu:
        one = cpu_baseline(ax, min(args.cpu_window, L), threads=1)
        allc = cpu_baseline(ax, L, threads=cpu_cores())
        allc["single_core"] = one
        out["cpu_baseline"] = allc
    m.close()
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
