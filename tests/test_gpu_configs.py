"""GPU parity at the benchmark configurations' full sizes.

c1  N=5000 x L=1500 rRNA-like alignment: int8 counts bit-exact against the oracle over all
    1,124,250 pairs; GT/MI/MIr/MIg/CHI (+APC, GT also +ASC) in FP64 within 1e-10 and in FP32
    within 1e-5; position-based weights on the device against the oracle's, split-BF16 weighted
    counts and every statistic against the oracle's FP64 weighted path, errors reported.
c4  N=20000 x L=5000 (1200 nested + 40 pseudoknotted pairs): counts bit-exact on three
    300-column windows at full N, row-sum invariants on 200 sampled columns against every other
    column; GT/MI/MIr/MIg/CHI each with APC and ASC against the oracle over all 12.5 million
    pairs; RAF/RAFS from counts on the windows and their APC/ASC over the whole matrix; weighted
    counts on a window.
c2  1000 simulated null alignments of uneven size (L uniform 80-600, N log-uniform 50-4000,
    mean branch length 0.05-0.3), GTp, one batch: every fit finite, 20 random alignments
    against the single-alignment path (which the oracle pins), graph replay = eager.

The relative error is floored at 1e-3 of the score span near zero (as in test_gpu_parity).
"""
import numpy as np
import pytest

import oracle
import rscape_amd as R
from rscape_amd import synth

pytestmark = pytest.mark.gpu


def _relerr(got, ref):
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    floor = max(1e-3 * (ref.max() - ref.min()), 1e-300)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor)))


def _upper_mask(L):
    return np.triu(np.ones((L, L), dtype=bool), 1)


# ---------------------------------------------------------------------------------- c1
@pytest.fixture(scope="module")
def c1():
    print("generating c1 (N=5000 x L=1500)", flush=True)
    return synth.config(1, seed=42)["ax"]


@pytest.fixture(scope="module")
def c1_oracle(c1):
    print("oracle probabilities at c1", flush=True)
    return oracle.probs(c1)


def test_c1_counts_bit_exact(ctx, c1):
    ax = c1
    N, L = ax.shape
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    got = m.counts().reshape(L, L, 16)
    m.close()
    ref = oracle.pair_counts(ax).reshape(L, L, 16)
    up = _upper_mask(L)
    bad = int(np.count_nonzero(np.any(got[up] != ref[up], axis=1)))
    print("c1 counts: %d of %d pairs differ" % (bad, up.sum()))
    assert bad == 0


@pytest.mark.parametrize("name,otype", [("GT", oracle.GT), ("MI", oracle.MI), ("MIr", oracle.MIr),
                                        ("MIg", oracle.MIg), ("CHI", oracle.CHI)])
def test_c1_statistics_apc(ctx, c1, c1_oracle, name, otype):
    ax = c1
    L = ax.shape[1]
    ref, _ = oracle.cov(c1_oracle, otype, oracle.CSELECT)
    corrs = [("p", oracle.APC)] + ([("a", oracle.ASC)] if name == "GT" else [])
    m = ctx.msa(ax)
    up = _upper_mask(L)
    for suffix, act in corrs:
        refc, mm = oracle.corrected(ref, L, act)
        refc = refc.reshape(L, L)
        for prec, bound in ((R.STAT_FP64, 1e-10), (R.STAT_FP32, 1e-5)):
            mn, mx = m.calculate(name + suffix, prec=prec)
            got = m.get_cov()[0]
            e = _relerr(got[up], refc[up])
            print("c1 %s%s %s: max rel err %.3e" % (name, suffix, "fp64" if prec else "fp32", e), flush=True)
            assert e < bound
            assert np.allclose([mn, mx], mm, rtol=10 * bound, atol=1e-9)
    m.close()


def test_c1_weighted(ctx, c1):
    """R-scape weights c1 with position-based weights (N > 1000, src/R-scape.c:1554-1557): the
    device weights against the oracle's, the split-BF16 counts (hi+lo and hi+lo+lo2) against the
    oracle's FP64 weighted counts, and each statistic (+APC) from BF16X3 counts with FP32
    statistics against the oracle's FP64 weighted path."""
    ax = c1
    N, L = ax.shape
    w = ctx.weights(ax, "PB")
    wr = oracle.weights_pb(ax)
    assert np.allclose(w, wr, rtol=1e-12, atol=0)
    print("oracle weighted probabilities at c1", flush=True)
    pr = oracle.probs(ax, w)
    up = _upper_mask(L)
    pp = pr["pp"].reshape(L, L, 16)[up]
    ns = pr["nseff"].reshape(L, L)[up]
    ref_cnt = pp * (ns[:, None] + 16e-10) - 1e-10
    m = ctx.msa(ax, w)
    # FP32 accumulation over N = 5000 weights bounds both splits (the third BF16 part removes the
    # weight-splitting error, not the accumulation's)
    for mode, bound in ((R.COUNTS_BF16, 1e-5), (R.COUNTS_BF16X3, 5e-6)):
        m.probs(mode)
        g = m.counts().reshape(L, L, 16)[up]
        rel = float(np.max(np.abs(g - ref_cnt) / np.maximum(np.abs(ref_cnt), 1.0)))
        print("c1 weighted counts mode %d: max rel err %.3e" % (mode, rel), flush=True)
        assert rel < bound
    # int8 weight limbs: only the 21-bit weight rounding (2^-22 max w per sequence) and one FP32
    # rounding of the record
    m.probs(R.COUNTS_W8)
    g = m.counts().reshape(L, L, 16)[up]
    rel8 = float(np.max(np.abs(g - ref_cnt) / np.maximum(np.abs(ref_cnt), 1.0)))
    print("c1 weighted counts int8 limbs: max rel err %.3e" % rel8, flush=True)
    assert rel8 < 2.0 ** -22 * w.max() / w.min() + 2.0 ** -22
    del ref_cnt, pp, g
    for name, otype in (("GT", oracle.GT), ("MI", oracle.MI), ("MIr", oracle.MIr), ("MIg", oracle.MIg),
                        ("CHI", oracle.CHI)):
        ref, _ = oracle.cov(pr, otype, oracle.CSELECT)
        refc, mm = oracle.corrected(ref, L, oracle.APC)
        refc = refc.reshape(L, L)
        m.calculate(name + "p", counts=R.COUNTS_BF16X3, prec=R.STAT_FP32)
        e = _relerr(m.get_cov()[0][up], refc[up])
        print("c1 weighted %sp (BF16X3 counts, FP32 statistic) vs FP64: max rel err %.3e" % (name, e), flush=True)
        # the FP32-accumulated counts carry ~2e-6; MIg subtracts ngap / nseff with ngap formed as
        # sum w - nseff, which magnifies the count error where few sequences are gapped
        assert e < (1e-3 if name == "MIg" else 1e-4)
        m.calculate(name + "p", counts=R.COUNTS_W8, prec=R.STAT_FP32)
        e8 = _relerr(m.get_cov()[0][up], refc[up])
        print("c1 weighted %sp (int8-limb counts, FP32 statistic) vs FP64: max rel err %.3e" % (name, e8), flush=True)
        assert e8 < 1e-5
    m.close()


# ---------------------------------------------------------------------------------- c4
WINDOWS = [(0, 300), (2350, 2650), (4700, 5000)]


@pytest.fixture(scope="module")
def c4():
    print("generating c4 (N=20000 x L=5000)", flush=True)
    return synth.config(4, seed=42)


@pytest.fixture(scope="module")
def c4_counts(ctx, c4):
    ax = c4["ax"]
    L = ax.shape[1]
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    cnt = m.counts().reshape(L, L, 16)
    m.close()
    return cnt


def test_c4_planted_pseudoknots(c4):
    ct = c4["ct"]
    L = len(ct)
    assert len(c4["pk"]) == 40 and int((ct >= 0).sum()) // 2 == 1240
    ks = np.nonzero(ct > np.arange(L))[0]
    ls = ct[ks]
    crossing = [np.any(((i < ks) & (ks < j) & (j < ls)) | ((ks < i) & (i < ls) & (ls < j))) for i, j in c4["pk"]]
    assert sum(crossing) >= 20


def test_c4_counts_windows_and_invariants(c4, c4_counts):
    ax = c4["ax"]
    N, L = ax.shape
    cnt = c4_counts
    for a, b in WINDOWS:
        print("c4 window %d..%d" % (a, b), flush=True)
        ref = oracle.pair_counts(np.ascontiguousarray(ax[:, a:b])).reshape(b - a, b - a, 16)
        got = cnt[a:b, a:b]
        up = _upper_mask(b - a)
        assert np.array_equal(got[up], ref[up].astype(float))
    # row sums: sum_y c_ij[x][y] = #{s: a_si = x, a_sj canonical} for 200 sampled i, all j
    rng = np.random.default_rng(3)
    cols = np.sort(rng.choice(L, 200, replace=False))
    canon = (ax < 4).astype(np.float32)
    for x in range(4):
        A = (ax[:, cols] == x).astype(np.float32)
        Rx = A.T @ canon                                   # exact: integers below 2^24
        c4x = cnt.reshape(L, L, 4, 4)
        for k, i in enumerate(cols):
            hi = np.arange(i + 1, L)                        # records (i, j), i < j: row sums
            lo = np.arange(0, i)                            # records (j, i), j < i: column sums
            assert np.array_equal(c4x[i, hi, x, :].sum(1), Rx[k, hi])
            assert np.array_equal(c4x[lo, i, :, x].sum(1), Rx[k, lo])
    assert np.all(cnt[np.tril(np.ones((L, L), dtype=bool))] == 0)


@pytest.fixture(scope="module")
def c4_oracle(c4):
    print("oracle probabilities at c4 (threads %d)" % oracle.lib().ref_get_threads(), flush=True)
    return oracle.probs(c4["ax"])


@pytest.mark.parametrize("name,otype", [("GT", oracle.GT), ("MI", oracle.MI), ("MIr", oracle.MIr),
                                        ("MIg", oracle.MIg), ("CHI", oracle.CHI)])
def test_c4_statistics_apc_asc(ctx, c4, c4_oracle, name, otype):
    ax = c4["ax"]
    L = ax.shape[1]
    ref, _ = oracle.cov(c4_oracle, otype, oracle.CSELECT)
    m = ctx.msa(ax)
    up = _upper_mask(L)
    for suffix, act in (("p", oracle.APC), ("a", oracle.ASC)):
        refc, mm = oracle.corrected(ref, L, act)
        refc = refc.reshape(L, L)
        for prec, bound in ((R.STAT_FP64, 1e-10), (R.STAT_FP32, 1e-5)):
            mn, mx = m.calculate(name + suffix, prec=prec)
            e = _relerr(m.get_cov()[0][up], refc[up])
            print("c4 %s%s %s: max rel err %.3e" % (name, suffix, "fp64" if prec else "fp32", e), flush=True)
            assert e < bound
            assert np.allclose([mn, mx], mm, rtol=10 * bound, atol=1e-9)
    m.close()


def test_c4_raf_rafs(ctx, c4):
    """RAF / RAFS depend on columns i and j only (RAFS also on the two anti-diagonal neighbours),
    so the oracle's count-based RAF of a window pins the device's values on that window; the
    APC / ASC corrections are then checked over the whole matrix from the device's uncorrected
    scores."""
    ax = c4["ax"]
    N, L = ax.shape
    m = ctx.msa(ax)
    for name in ("RAF", "RAFS"):
        m.calculate(name)
        raw = m.get_cov()[0]
        for a, b in WINDOWS:
            sub = np.ascontiguousarray(ax[:, a:b])
            n = b - a
            ref, _ = oracle.raf_counts(oracle.pair_counts(sub), N, n)
            if name == "RAFS":
                ref, _ = oracle.rafs_from(ref, n)
            ref = ref.reshape(n, n)
            up = _upper_mask(n)
            if name == "RAFS":                 # the window's edge pairs need neighbours outside it
                up[0, :] = False; up[:, n - 1] = False
            assert np.allclose(raw[a:b, a:b][up], ref[up], rtol=1e-12, atol=1e-12), (name, a)
        for suffix, act in (("p", oracle.APC), ("a", oracle.ASC)):
            refc, mm = oracle.corrected(raw.ravel().copy(), L, act)
            mn, mx = m.calculate(name + suffix)
            e = _relerr(m.get_cov()[0][_upper_mask(L)], refc.reshape(L, L)[_upper_mask(L)])
            print("c4 %s%s: max rel err %.3e" % (name, suffix, e), flush=True)
            assert e < 1e-10
    m.close()


def test_c4_weighted_window(ctx, c4):
    """Position-based weights and split-BF16 weighted counts at full N on a 300-column window."""
    ax = c4["ax"]
    a, b = WINDOWS[1]
    sub = np.ascontiguousarray(ax[:, a:b])
    w = ctx.weights(ax, "PB")
    assert np.allclose(w, oracle.weights_pb(ax), rtol=1e-12, atol=0)
    pr = oracle.probs(sub, w)
    n = b - a
    up = _upper_mask(n)
    ref = pr["pp"].reshape(n, n, 16)[up] * (pr["nseff"].reshape(n, n)[up][:, None] + 16e-10) - 1e-10
    m = ctx.msa(ax, w)
    m.probs(R.COUNTS_BF16X3)
    got = m.counts().reshape(ax.shape[1], ax.shape[1], 16)[a:b, a:b][up]
    m.close()
    rel = float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)))
    print("c4 weighted counts (BF16X3) on columns %d..%d: max rel err %.3e" % (a, b, rel))
    assert rel < 5e-6


# ---------------------------------------------------------------------------------- c2
def c2_jobs(n=1000, seed=42):
    from rscape_amd import null as NL
    return NL.uneven_jobs(n, seed)


def test_c2_thousand_alignment_batch(ctx):
    """Every one of the 1000 batched nulls equals the single-alignment path (simulate, score,
    width, histogram, fit), including the alignments whose tail admits no finite fit."""
    from rscape_amd import null as NL
    jobs = c2_jobs()
    print("c2: %d jobs, %.3g updates" % (len(jobs), sum(j.tree["n"] * j.L * (j.L - 1) / 2 for j in jobs)), flush=True)
    fits, hs = NL.batch(ctx, jobs, "GTp", pmass=5e-4, hists=True)
    assert fits.shape == (1000, 8)
    assert all(h is not None for h in hs)
    nonfinite = 0
    for k, job in enumerate(jobs):
        msa = ctx.msa(NL.simulate(ctx, job))
        mn, mx = msa.calculate("GTp")
        wd = R.width(mx, mn)
        h = msa.rank_hist(-10.0, wd)
        fit, _ = h.fit(ctx, 5e-4, 1.0, True)
        msa.close()
        assert fits[k][5] == wd and np.isclose(fits[k][6], mn, rtol=1e-12) and np.isclose(fits[k][7], mx, rtol=1e-12), k
        assert np.array_equal(hs[k].h.counts(), h.h.counts()[:hs[k].h.nb]), k
        ref = np.array([fit[key] for key in ["mu", "lambda", "tau", "newmass", "pmass"]])
        assert np.allclose(fits[k][:5], ref, rtol=1e-9, atol=1e-12, equal_nan=True), (k, fits[k][:5], ref)
        nonfinite += int(not np.all(np.isfinite(fits[k][:5])))
    print("c2: %d of 1000 tails without a finite fit (same on the single path)" % nonfinite)
    assert nonfinite < 100
    g, _ = NL.batch(ctx, jobs, "GTp", pmass=5e-4, graph=2)
    assert np.array_equal(np.nan_to_num(g), np.nan_to_num(fits))
