"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Counts must be bit-exact; FP64 statistics within 1e-10 relative (with the floor
|ref| >= 1e-3 * (maxCOV - minCOV) used near zero); FP32 statistics within the
looser bound stated in each test.
"""
import numpy as np
import pytest

import oracle
import rscape_amd as R
from rscape_amd import synth

pytestmark = pytest.mark.gpu


def _msa(N, L, seed, gap=0.15, other=0.03):
    rng = np.random.default_rng(seed)
    return synth.random_msa(N, L, rng, gap=gap, other=other)


def _upper(a, L):
    a = np.asarray(a).reshape(L, L, -1) if np.asarray(a).size > L * L else np.asarray(a).reshape(L, L)
    iu = np.triu_indices(L, 1)
    return a[iu]


def _relerr(got, ref):
    got, ref = np.asarray(got, float), np.asarray(ref, float)
    span = (ref.max() - ref.min()) if ref.size else 1.0
    floor = max(1e-3 * span, 1e-300)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), floor))) if ref.size else 0.0


@pytest.mark.parametrize("N,L", [(1, 2), (7, 3), (64, 37), (65, 64), (200, 65), (333, 130), (129, 200)])
def test_int8_counts_bit_exact(ctx, N, L):
    print("counts case N=%d L=%d" % (N, L), flush=True)
    ax = _msa(N, L, 1000 + N + L)
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    got = m.counts().reshape(L, L, 16)
    ref = oracle.pair_counts(ax).reshape(L, L, 16)
    iu = np.triu_indices(L, 1)
    bad = np.nonzero(np.any(got[iu] != ref[iu], axis=1))[0]
    for k in bad[:4]:
        print("mismatch pair (%d,%d)\n got %s\n ref %s" % (iu[0][k], iu[1][k], got[iu][k].astype(int), ref[iu][k]))
    assert np.array_equal(got[iu], ref[iu].astype(float))


@pytest.mark.parametrize("shape", range(1, 10))
def test_gram_shapes_bit_exact(ctx, shape):
    print("gram shape %d (%s)" % (shape, R.lib().rsg_gram_shape_name(shape).decode()), flush=True)
    N, L = 300, 203
    ax = _msa(N, L, 7 + shape)
    ctx.set_gram_shape(shape)
    try:
        m = ctx.msa(ax)
        m.probs(R.COUNTS_INT8)
        got = m.counts().reshape(L, L, 16)
        pm = m.get_probs()["pm"]
    finally:
        ctx.set_gram_shape(0)
    ref = oracle.pair_counts(ax).reshape(L, L, 16)
    iu = np.triu_indices(L, 1)
    assert np.array_equal(got[iu], ref[iu].astype(float))
    # the marginals are summed per 32 x 32 column block in a fixed order whatever the tile shape
    m0 = ctx.msa(ax)
    m0.probs(R.COUNTS_INT8)
    assert np.array_equal(pm, m0.get_probs()["pm"])
    assert np.allclose(pm, oracle.probs(ax)["pm"], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("N", [257, 288, 289])
@pytest.mark.parametrize("shape", range(1, 10))
def test_gram_tail_step_bit_exact(ctx, shape, N):
    """N = 257 and 288 leave 1 and 32 sequences past the last whole K-tile of every shape (64,
    128, 256): the int8 Gram runs them as one k-step from registers after its loop; N = 289
    (33 past) takes a whole padded K-tile instead.  Counts and marginals as the oracle's."""
    print("gram tail: shape %d N=%d" % (shape, N), flush=True)
    L = 203
    ax = _msa(N, L, 31 + shape + N)
    ctx.set_gram_shape(shape)
    try:
        m = ctx.msa(ax)
        m.probs(R.COUNTS_INT8)
        got = m.counts().reshape(L, L, 16)
        pm = m.get_probs()["pm"]
    finally:
        ctx.set_gram_shape(0)
    ref = oracle.pair_counts(ax).reshape(L, L, 16)
    iu = np.triu_indices(L, 1)
    assert np.array_equal(got[iu], ref[iu].astype(float))
    assert np.allclose(pm, oracle.probs(ax)["pm"], rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("N,L", [(600, 1500), (300, 203), (5000, 90)])
def test_gram_split_matches_unsplit(ctx, N, L):
    """Tiles of the last partial wave split along K (partials summed) give the same counts."""
    print("gram split N=%d L=%d" % (N, L), flush=True)
    ax = _msa(N, L, 31 + L)
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    a, pa = m.counts(), m.get_probs()["pm"]
    ctx.set_gram_split(False)
    try:
        m.probs(R.COUNTS_INT8)
        b, pb = m.counts(), m.get_probs()["pm"]
    finally:
        ctx.set_gram_split(True)
    assert np.array_equal(a, b)
    # split tiles get their marginal partials from the fix-up pass, in the epilogue's order
    assert np.array_equal(pa, pb)
    ref = oracle.pair_counts(ax).reshape(L, L, 16)
    iu = np.triu_indices(L, 1)
    assert np.array_equal(a.reshape(L, L, 16)[iu], ref[iu].astype(float))


def test_counts_u32_path(ctx):
    """N >= 65536 switches the records to 32-bit counts."""
    N, L = 70000, 9
    print("u32 counts N=%d L=%d" % (N, L), flush=True)
    ax = _msa(N, L, 5)
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    got = m.counts().reshape(L, L, 16)
    ref = oracle.pair_counts(ax).reshape(L, L, 16)
    iu = np.triu_indices(L, 1)
    assert np.array_equal(got[iu], ref[iu].astype(float))


def test_probs_match_oracle(ctx):
    N, L = 157, 71
    ax = _msa(N, L, 11)
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    g = m.get_probs()
    r = oracle.probs(ax)
    assert np.allclose(g["pm"], r["pm"], rtol=1e-12, atol=1e-15)
    assert np.allclose(g["ps"], r["ps"], rtol=1e-12, atol=1e-15)
    assert np.allclose(g["pp"], r["pp"], rtol=1e-12, atol=1e-18)
    assert np.array_equal(g["nseff"], r["nseff"])
    iu = np.triu_indices(L, 1)
    assert np.allclose(g["ngap"].reshape(L, L)[iu], r["ngap"].reshape(L, L)[iu], rtol=0, atol=1e-9)


STATS = [("CHI", oracle.CHI), ("GT", oracle.GT), ("MI", oracle.MI), ("MIr", oracle.MIr),
         ("MIg", oracle.MIg), ("OMES", oracle.OMES), ("CCF", oracle.CCF)]


@pytest.mark.parametrize("name,otype", STATS)
@pytest.mark.parametrize("cls", [R.C16, R.C2])
def test_statistics_fp64(ctx, name, otype, cls):
    N, L = 90, 57
    print("stat %s class %d" % (name, cls), flush=True)
    ax = _msa(N, L, 21)
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    m.cov(name, cls, prec=R.STAT_FP64)
    got, mn, mx = m.get_cov()
    pr = oracle.probs(ax)
    ref, mm = oracle.cov(pr, otype, cls)
    assert _relerr(_upper(got, L), _upper(ref, L)) < 1e-10
    assert np.allclose([mn, mx], mm, rtol=1e-10, atol=1e-12)
    assert np.all(np.isneginf(np.diag(got)))


def test_gt_cwc_and_cselect(ctx):
    ax = _msa(6, 40, 3)            # nseq <= 8 -> CSELECT resolves to C2
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    pr = oracle.probs(ax)
    for cls in (R.CWC, R.CSELECT):
        m.cov("GT", cls)
        got, _, _ = m.get_cov()
        ref, _ = oracle.cov(pr, oracle.GT, cls)
        assert _relerr(_upper(got, 40), _upper(ref, 40)) < 1e-10


@pytest.mark.parametrize("actype", [R.APC, R.ASC])
def test_corrections(ctx, actype):
    N, L = 120, 77
    ax = _msa(N, L, 31)
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    m.cov("GT", R.C16)
    m.correct(actype)
    got, mn, mx = m.get_cov()
    pr = oracle.probs(ax)
    ref, _ = oracle.cov(pr, oracle.GT, R.C16)
    ref2, mm = oracle.corrected(ref, L, actype)
    assert _relerr(_upper(got, L), _upper(ref2, L)) < 1e-10
    off = ~np.eye(L, dtype=bool)
    assert _relerr(got[off], ref2.reshape(L, L)[off]) < 1e-10
    assert np.allclose([mn, mx], mm, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("rafs", [False, True])
def test_raf_rafs(ctx, rafs):
    N, L = 40, 33
    ax = _msa(N, L, 41)
    m = ctx.msa(ax)
    m.cov("RAFS" if rafs else "RAF")
    got, mn, mx = m.get_cov()
    ref, mm = oracle.raf(ax, rafs)
    off = ~np.eye(L, dtype=bool)
    assert np.allclose(got[off], ref.reshape(L, L)[off], rtol=1e-12, atol=1e-14)
    assert np.allclose([mn, mx], mm, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("N,L,cls", [(400, 60, R.C16), (400, 60, R.C2), (5000, 160, R.C16)])
def test_fp32_statistics_error_bound(ctx, N, L, cls):
    """FP32 statistics (FP64 differences, FP32 logarithms): within 1e-5 of the FP64 oracle
    (relative, with the near-zero floor), every statistic."""
    print("fp32 stats N=%d L=%d class %d" % (N, L, cls), flush=True)
    ax = _msa(N, L, 51 + N)
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    pr = oracle.probs(ax)
    for name, ot in STATS[:6]:
        m.cov(name, cls, prec=R.STAT_FP32)
        got, mn, mx = m.get_cov()
        ref, mm = oracle.cov(pr, ot, cls)
        err = _relerr(_upper(got, L), _upper(ref, L))
        print("fp32 %s max rel err (floored) %.3e" % (name, err))
        assert err < 1e-5, name
        assert np.all(np.isneginf(np.diag(got)))


def test_fp32_weighted_statistics(ctx):
    """FP32 statistics on FP32 (split-BF16) weighted counts against FP64 statistics on the same
    counts: the statistic adds at most 1e-5 to the count error."""
    N, L = 700, 96
    rng = np.random.default_rng(71)
    ax = synth.random_msa(N, L, rng, gap=0.1)
    w = rng.uniform(0.05, 3.0, size=N)
    m = ctx.msa(ax, w)
    m.probs(R.COUNTS_BF16X3)
    for name in ("GT", "MI", "MIr", "MIg", "CHI", "OMES"):
        m.cov(name, R.C16, prec=R.STAT_FP64)
        ref, _, _ = m.get_cov()
        m.cov(name, R.C16, prec=R.STAT_FP32)
        got, _, _ = m.get_cov()
        err = _relerr(_upper(got, L), _upper(ref, L))
        print("weighted fp32 %s vs fp64 on the same counts: %.3e" % (name, err))
        assert err < 1e-5, name


@pytest.mark.parametrize("mode,bound", [(R.COUNTS_BF16, 2e-5), (R.COUNTS_BF16X3, 1e-6)])
def test_weighted_counts_split_bf16(ctx, mode, bound):
    N, L = 500, 48
    rng = np.random.default_rng(61)
    ax = synth.random_msa(N, L, rng, gap=0.1)
    w = rng.uniform(0.05, 3.0, size=N)
    m = ctx.msa(ax, w)
    m.probs(mode)
    got = m.counts().reshape(L, L, 16)
    pr = oracle.probs(ax, w)
    iu = np.triu_indices(L, 1)
    # FP64 weighted counts from the oracle probabilities
    pp = pr["pp"].reshape(L, L, 16)[iu]
    ns = pr["nseff"].reshape(L, L)[iu]
    ref = pp * (ns[:, None] + 16e-10) - 1e-10
    g = got[iu]
    rel = np.max(np.abs(g - ref) / np.maximum(np.abs(ref), 1.0))
    print("split-bf16 mode %d: max rel count error %.3e" % (mode, rel))
    assert rel < bound
    m.cov("GT", R.C16)
    gt, _, _ = m.get_cov()
    ref_gt, _ = oracle.cov(pr, oracle.GT, R.C16)
    e = _relerr(_upper(gt, L), _upper(ref_gt, L))
    print("split-bf16 mode %d: GT max rel err (floored) %.3e" % (mode, e))
    assert e < 50 * bound


def _weighted_counts_numpy(ax, w):
    """Textbook weighted pair counts: X^T diag(w) X over the one-hot matrix X (N x 4L) in
    float64; cnt[i, j, x*4+y] = sum_s w_s [a_si = x][a_sj = y]."""
    N, L = ax.shape
    X = np.zeros((N, L, 4))
    for x in range(4):
        X[:, :, x] = ax == x
    X = X.reshape(N, 4 * L)
    G = (X * w[:, None]).T @ X
    return G.reshape(L, 4, L, 4).transpose(0, 2, 1, 3).reshape(L, L, 16)


@pytest.mark.parametrize("spread", [(0.05, 3.0), (1e-3, 10.0)])
def test_weighted_counts_int8_limbs(ctx, spread):
    """COUNTS_W8: the weights rounded to 21-bit fixed point (q = rint(w / W), W = max w /
    (2^21 - 1)), three 7-bit limbs through the int8 Gram.  Against numpy's FP64 counts with the
    rounded weights the records are W * (an exact integer) rounded once to FP32 (<= 1 ulp);
    against the true weights the difference is at most the rounding, 2^-22 max w per sequence.
    Statistics from these counts are within 1e-5 of the FP64 oracle on the true weights."""
    print("w8 counts, weights U%s" % (spread,), flush=True)
    N, L = 700, 96
    rng = np.random.default_rng(67)
    ax = _msa(N, L, 67)
    w = np.exp(rng.uniform(np.log(spread[0]), np.log(spread[1]), size=N))
    m = ctx.msa(ax, w)
    m.probs(R.COUNTS_W8)
    got = m.counts().reshape(L, L, 16)
    iu = np.triu_indices(L, 1)
    g = got[iu]
    W = w.max() / (2 ** 21 - 1)
    wq = W * np.rint(w / W)
    refq = _weighted_counts_numpy(ax, wq)[iu]
    assert np.all(np.abs(g - refq) <= 2.0 ** -23 * np.abs(refq) + 1e-300)
    ref = _weighted_counts_numpy(ax, w)[iu]
    nunit = _weighted_counts_numpy(ax, np.ones(N))[iu]
    assert np.all(np.abs(g - ref) <= 2.0 ** -22 * w.max() * nunit + 2.0 ** -23 * np.abs(ref))
    print("w8 counts: max rel err vs true weights %.3e" % np.max(np.abs(g - ref) / np.maximum(ref, 1.0)))
    pr = oracle.probs(ax, w)
    for name, ot in STATS[:6]:
        m.cov(name, R.C16, prec=R.STAT_FP32)
        got_s, _, _ = m.get_cov()
        ref_s, _ = oracle.cov(pr, ot, R.C16)
        e = _relerr(_upper(got_s, L), _upper(ref_s, L))
        print("w8 counts + FP32 %s vs FP64 weighted oracle: %.3e" % (name, e))
        assert e < 1e-5, name


def test_weighted_counts_int8_limbs_split_matches_unsplit(ctx):
    """COUNTS_W8 with the tiles of the last partial wave split along K (the exact int64 partial
    sums added by k_gram_w8_fix) equals the unsplit kernel bit for bit: the records and the exact
    nseff (seen through MIg).  L = 1100 gives 630 tiles of 32 x 32 columns, 118 past the 512
    work-group slots of 256 CUs, so the tail is split."""
    N, L = 300, 1100
    print("w8 split N=%d L=%d" % (N, L), flush=True)
    rng = np.random.default_rng(73)
    ax = _msa(N, L, 73)
    w = rng.uniform(0.05, 3.0, size=N)
    m = ctx.msa(ax, w)
    m.probs(R.COUNTS_W8)
    a = m.counts()
    m.cov("MIg", R.C16, prec=R.STAT_FP64)
    ga = m.get_cov()[0]
    ctx.set_gram_split(False)
    try:
        m.probs(R.COUNTS_W8)
        b = m.counts()
        m.cov("MIg", R.C16, prec=R.STAT_FP64)
        gb = m.get_cov()[0]
    finally:
        ctx.set_gram_split(True)
    assert np.array_equal(a, b)
    assert np.array_equal(np.nan_to_num(ga), np.nan_to_num(gb))


PIPE = [("GTp", oracle.GT, R.APC), ("GT", oracle.GT, None), ("GTa", oracle.GT, R.ASC),
        ("MIp", oracle.MI, R.APC), ("CCFp", oracle.CCF, R.APC), ("RAF", oracle.RAF, None),
        ("RAFS", oracle.RAFS, None), ("CHIa", oracle.CHI, R.ASC)]


@pytest.mark.parametrize("covtype,otype,corr", PIPE)
@pytest.mark.parametrize("wait", [2000, 0])
def test_full_pipeline_calculate(ctx, covtype, otype, corr, wait):
    """rsg_cov_Calculate (one call: counts, probabilities, statistic, correction, extremes) for
    every way its chain ends -- the correction's reduction writing the result block (APC, ASC),
    the statistic's own extremes (no correction), RAF / RAFS (their min / max pass), CCF (a host
    step mid-chain) -- with the completion stamp waited for, and with the stamp wait off (the
    stream synchronisation fallback)."""
    print("pipeline %s wait %d" % (covtype, wait), flush=True)
    N, L = 200, 120
    d = synth.config(0, seed=3)
    ax = d["ax"]
    ctx.set_stamp_wait(wait)
    try:
        m = ctx.msa(ax)
        mn, mx = m.calculate(covtype)
        got, gmn, gmx = m.get_cov()
    finally:
        ctx.set_stamp_wait(2000)
    if otype in (oracle.RAF, oracle.RAFS):
        ref, mm = oracle.raf(ax, otype == oracle.RAFS)
    else:
        pr = oracle.probs(ax)
        ref, mm = oracle.cov(pr, otype, R.CSELECT)
    if corr is not None:
        ref, mm = oracle.corrected(ref, L, corr)
    assert _relerr(_upper(got, L), _upper(ref, L)) < 1e-10
    assert np.allclose([mn, mx], mm, rtol=1e-10, atol=1e-12)
    assert (gmn, gmx) == (mn, mx)
    assert np.all(np.isneginf(np.diag(got)))


def test_tutorial_golden_gtp(ctx):
    """End-to-end against the published tutorial scores (GSC weights, gap filter, GT, APC)."""
    import json, os
    from rscape_amd import stockholm
    here = os.path.join(os.path.dirname(__file__), "golden")
    names, ax, _ = stockholm.read(os.path.join(here, "updated_Arisong.sto"))
    gold = json.load(open(os.path.join(here, "arisong_gtp.json")))
    use = oracle.gapcol_filter(ax, np.ones(len(names)), 0.75)
    cols = np.nonzero(use)[0]
    ax2 = oracle.canonicalize(ax[:, cols])
    w = oracle.weights_gsc(ax2)
    m = ctx.msa(ax2, w)
    m.calculate("GTp", counts=R.COUNTS_BF16X3)
    got, _, mx = m.get_cov()
    pos = {c + 1: k for k, c in enumerate(cols)}
    errs = [abs(got[pos[i], pos[j]] - sc) / sc for i, j, sc in gold["pairs"]]
    assert max(errs) < 1e-3 and np.median(errs) < 5e-5, errs    # as the oracle (test_oracle.py)
    # and the device path against the oracle on the same weights (BF16X3 counts, FP64 statistics)
    pr = oracle.probs(ax2, w)
    ref, _ = oracle.cov(pr, oracle.GT, oracle.CSELECT)
    ref, _ = oracle.corrected(ref, len(cols), oracle.APC)
    assert _relerr(_upper(got, len(cols)), _upper(ref, len(cols))) < 1e-4


@pytest.mark.parametrize("N,L", [(40, 2), (40, 33), (150, 65), (64, 97), (300, 161)])
def test_marginals_and_gtp_edge_columns(ctx, N, L):
    """Marginals from the Gram's 32 x 32 block partials at sizes whose last block holds one
    column (L = 32 b + 1) or two, with an all-gap column and an all-non-canonical column (their
    pair sums are 0: pm falls back to 1/4 as esl_vec_DNorm does), then the whole GTp scoring."""
    print("edge columns N=%d L=%d" % (N, L), flush=True)
    ax = _msa(N, L, 100 + L)
    ax[:, 0] = 4                                  # all gaps
    ax[:, L - 1] = 15                             # all non-canonical
    m = ctx.msa(ax)
    m.probs(R.COUNTS_INT8)
    g = m.get_probs()
    pr = oracle.probs(ax)
    assert np.allclose(g["pm"], pr["pm"], rtol=1e-12, atol=1e-15)
    assert np.allclose(g["pm"].reshape(L, 4)[0], 0.25) and np.allclose(g["pm"].reshape(L, 4)[L - 1], 0.25)
    mn, mx = m.calculate("GTp")
    got, _, _ = m.get_cov()
    ref, _ = oracle.cov(pr, oracle.GT, R.CSELECT)
    ref2, mm = oracle.corrected(ref, L, R.APC)
    assert _relerr(_upper(got, L), _upper(ref2, L)) < 1e-10
    assert np.allclose([mn, mx], mm, rtol=1e-10, atol=1e-12)
