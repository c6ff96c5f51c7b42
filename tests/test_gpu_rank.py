"""GPU parity of ranking, null-histogram merge, tail fits and E-values (FP64)."""
import os

import numpy as np
import pytest

import oracle
import rscape_amd as R
from rscape_amd import synth

pytestmark = pytest.mark.gpu


def scored(ctx, ax, covtype="GTp"):
    m = ctx.msa(ax)
    mn, mx = m.calculate(covtype)
    cov, _, _ = m.get_cov()
    return m, cov, mn, mx


def same_hist(g, r):
    g, r = g.h, oracle._hp(r).contents
    assert (g.nb, g.imin, g.imax, g.n, g.Nc, g.No) == (r.nb, r.imin, r.imax, r.n, r.Nc, r.No)
    assert g.bmin == r.bmin and g.bmax == r.bmax and g.w == r.w
    assert g.xmin == r.xmin and g.xmax == r.xmax
    assert np.array_equal(g.counts(), r.counts())


def test_rank_histogram_matches_oracle(ctx):
    ax = synth.config(0, seed=5)["ax"]
    L = ax.shape[1]
    m, cov, mn, mx = scored(ctx, ax)
    w = R.width(mx, mn)
    g = m.rank_hist(-10.0, w)
    r = oracle.rank_hist(cov, L, mx, -10.0, w)
    same_hist(g, r)


def null_replicates(ctx, n=4, seed=0):
    """Score unstructured replicate alignments; return (gpu histograms, oracle histograms)."""
    rng = np.random.default_rng(seed)
    gh, rh, w = [], [], None
    for k in range(n):
        ax = synth.random_msa(60, 50 + 7 * k, rng, gap=0.1)
        m, cov, mn, mx = scored(ctx, ax)
        if w is None:
            w = R.width(mx, mn)
        gh.append(m.rank_hist(-10.0, w))
        rh.append(oracle.rank_hist(cov, ax.shape[1], mx, -10.0, w))
    return gh, rh


def test_null_merge_matches_oracle(ctx):
    gh, rh = null_replicates(ctx)
    gcum, rcum = None, None
    for g, r in zip(gh, rh):
        gcum = R.Histogram.null_add(gcum, g)
        rcum = oracle.null_add(rcum, r)
    same_hist(gcum, rcum)


@pytest.mark.parametrize("expo", [True, False])
@pytest.mark.parametrize("pmass", [5e-4, 0.05])
def test_null_fit_matches_oracle(ctx, expo, pmass):
    gh, rh = null_replicates(ctx, n=6, seed=1)
    gcum, rcum = None, None
    for g, r in zip(gh, rh):
        gcum = R.Histogram.null_add(gcum, g)
        rcum = oracle.null_add(rcum, r)
    gf, gsf = gcum.fit(ctx, pmass, 1.0, expo)
    rf, rsf = oracle.null_fit(rcum, pmass, 1.0, expo)
    print("fit gpu", gf, "oracle", rf)
    for k in ("mu", "lambda", "tau", "newmass", "pmass"):
        assert np.isclose(gf[k], rf[k], rtol=1e-10, atol=1e-12), k
    assert (gsf is None) == (rsf is None)
    if gsf is not None:
        assert np.allclose(gsf, rsf, rtol=1e-10, atol=1e-300)
    rr = oracle._hp(rcum).contents
    assert (gcum.h.cmin, gcum.h.No, gcum.h.phi) == (rr.cmin, rr.No, rr.phi)


def test_evalues_match_oracle(ctx):
    gh, rh = null_replicates(ctx, n=6, seed=2)
    gcum, rcum = None, None
    for g, r in zip(gh, rh):
        gcum = R.Histogram.null_add(gcum, g)
        rcum = oracle.null_add(rcum, r)
    gf, gsf = gcum.fit(ctx, 0.05, 1.0, True)
    rf, rsf = oracle.null_fit(rcum, 0.05, 1.0, True)
    ax = synth.config(0, seed=9, N=60, L=70)["ax"]
    m, cov, mn, mx = scored(ctx, ax)
    L = ax.shape[1]
    Nc = L * (L - 1) // 2
    ge = m.evalues(gcum, gsf, Nc)
    re = oracle.evalues(cov, L, Nc, rcum, rsf)
    off = ~np.eye(L, dtype=bool)
    assert np.allclose(ge[off], re[off], rtol=1e-10, atol=0)
    # scalar helpers agree with the matrix
    i, j = np.unravel_index(np.argmax(np.where(off, cov, -np.inf)), cov.shape)
    assert np.isclose(gcum.cov2evalue(cov[i, j], Nc, gsf), ge[i, j], rtol=1e-12)
    # evalue2cov (covariation.c:2404-2436) against the oracle's restatement, several thresholds
    import ctypes as C
    L_ = oracle.lib()
    L_.ref_evalue2cov.argtypes = [C.c_double, C.c_int64, C.c_void_p, C.c_void_p]
    for e in (1e-6, 1e-3, 0.05, 1.0, 10.0, 1e4):
        thr = gcum.evalue2cov(e, Nc, gsf)
        ref = L_.ref_evalue2cov(e, Nc, rcum, rsf.ctypes.data_as(C.c_void_p))
        assert (thr == ref) or np.isclose(thr, ref, rtol=1e-12), (e, thr, ref)
    assert np.isfinite(gcum.evalue2cov(0.05, Nc, gsf))


def test_null_file_round_trip(ctx, tmp_path):
    gh, _ = null_replicates(ctx, n=3, seed=3)
    cum = None
    for g in gh:
        cum = R.Histogram.null_add(cum, g)
    path = str(tmp_path / "x.null")
    cum.write_null(path)
    lines = open(path).read().splitlines()
    assert len(lines) == cum.h.nb
    back = R.Histogram.read_null(path)
    assert back.h.n == cum.h.n and back.h.Nc == cum.h.Nc
    a = cum.h.counts(); b = back.h.counts()
    assert np.array_equal(a[a > 0], b[b > 0])          # same nonzero counts, in order


def test_rank_hist_rejects_empty_range(ctx):
    """bmin above every score leaves the histogram without bins: a clear error before any
    allocation or launch, and the context keeps working."""
    ax = synth.config(0, seed=5)["ax"]
    m, cov, mn, mx = scored(ctx, ax)
    with pytest.raises(R.RscapeError, match="empty|no bins"):
        m.rank_hist(mx + 100.0, 0.05)
    g = m.rank_hist(-10.0, R.width(mx, mn))
    assert int(g.h.counts().sum()) == ax.shape[1] * (ax.shape[1] - 1) // 2


def test_pm_validation_failure_is_reported(ctx):
    """corr_ValidateProbs (correlators.c:1499-1545) on the device: a negative tolerance makes every
    marginal fail validation; the device flag comes back as RSG_FAIL with a message from both
    rsg_corr_Probs and rsg_cov_Calculate, and the context keeps working.  Weights that are
    negative or not finite fail as pp validation."""
    ax = synth.config(0, seed=5)["ax"]
    m = ctx.msa(ax)
    with pytest.raises(R.RscapeError, match="pm validation"):
        m.probs(R.COUNTS_INT8, tol=-1.0)
    with pytest.raises(R.RscapeError, match="pm validation"):
        m.calculate("GTp", tol=-1.0)
    for bad in (np.nan, -1.0, np.inf):                       # pp cannot validate with these weights
        w = np.ones(ax.shape[0]); w[3] = bad
        mw = ctx.msa(ax, w)
        with pytest.raises(R.RscapeError, match="pp validation"):
            mw.calculate("GTp", counts=R.COUNTS_BF16)
    mn, mx = m.calculate("GTp")
    assert np.isfinite(mn) and np.isfinite(mx)
