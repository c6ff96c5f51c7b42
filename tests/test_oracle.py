"""CPU tests that pin the oracle (no GPU).

Each check is independent of the oracle's own formulas: brute-force counting, textbook
closed forms (G-test, Pearson chi-square via scipy), hand-computed corrections, the
published tutorial scores, scipy fits, Philox known-answer vectors and finite
differences.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest
from scipy import optimize, stats

import oracle
from oracle import P
from rscape_amd import synth, stockholm

HERE = os.path.dirname(os.path.abspath(__file__))


def brute_counts(ax, w=None):
    N, L = ax.shape
    w = np.ones(N) if w is None else w
    c = np.zeros((L, L, 4, 4))
    for s in range(N):
        for i in range(L):
            for j in range(i + 1, L):
                a, b = ax[s, i], ax[s, j]
                if a < 4 and b < 4:
                    c[i, j, a, b] += w[s]
    return c


def test_counts_brute_force_and_invariants():
    rng = np.random.default_rng(0)
    ax = synth.random_msa(23, 9, rng, gap=0.2, other=0.1)
    w = rng.uniform(0.1, 2, 23)
    ref = brute_counts(ax)
    got = oracle.pair_counts(ax).reshape(9, 9, 4, 4)
    iu = np.triu_indices(9, 1)
    assert np.array_equal(got[iu], ref[iu])
    pr = oracle.probs(ax, w)
    cw = brute_counts(ax, w)
    ns = pr["nseff"].reshape(9, 9)
    ng = pr["ngap"].reshape(9, 9)
    pp = pr["pp"].reshape(9, 9, 4, 4)
    for i, j in zip(*iu):
        assert np.isclose(ns[i, j], cw[i, j].sum())
        assert np.isclose(ns[i, j] + ng[i, j], w.sum())            # nseff + ngap = sum w
        assert np.allclose(pp[j, i], pp[i, j].T)                     # symmetrised
        assert np.allclose(pp[i, j], (cw[i, j] + 1e-10) / (cw[i, j].sum() + 16e-10))
        assert ng[j, i] == 0.0                                       # ngap[j][i] never written


def test_marginals_closed_form():
    rng = np.random.default_rng(1)
    ax = synth.random_msa(31, 6, rng, gap=0.3)
    pr = oracle.probs(ax)
    c = brute_counts(ax)
    L = 6
    for i in range(L):
        acc = np.zeros(4)
        for j in range(L):
            if j == i:
                continue
            t = c[min(i, j), max(i, j)]
            t = t if i < j else t.T
            if t.sum() > 0:
                acc += (t.sum(1) + 4e-10) / (t.sum() + 16e-10)
        assert np.allclose(pr["pm"][i * 4:(i + 1) * 4], acc / acc.sum(), rtol=1e-13)
    # ps includes the gap symbol with a 1e-5 prior
    for i in range(L):
        cnt = np.array([(ax[:, i] == x).sum() for x in range(5)], float) + 1e-5
        assert np.allclose(pr["ps"][i * 5:(i + 1) * 5], cnt / cnt.sum())


def two_column(table):
    """Alignment of 2 columns realising an integer 4x4 count table."""
    rows = []
    for x in range(4):
        for y in range(4):
            rows += [(x, y)] * int(table[x, y])
    return np.array(rows, dtype=np.uint8)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gt_chi_are_textbook_tests_at_L2(seed):
    rng = np.random.default_rng(seed)
    table = rng.integers(1, 30, size=(4, 4))
    ax = two_column(table)
    N = ax.shape[0]
    pr = oracle.probs(ax)
    gt, _ = oracle.cov(pr, oracle.GT, oracle.C16, nseqthresh=0, alenthresh=0)
    chi, _ = oracle.cov(pr, oracle.CHI, oracle.C16, nseqthresh=0, alenthresh=0)
    g_ref = stats.chi2_contingency(table, correction=False, lambda_="log-likelihood")[0]
    c_ref = stats.chi2_contingency(table, correction=False)[0]
    assert np.isclose(gt[1], g_ref, rtol=1e-7)
    assert np.isclose(chi[1], c_ref, rtol=1e-7)
    # MI = G / (2 N) (natural log) and OMES = chi-like sum / N
    mi, _ = oracle.cov(pr, oracle.MI, oracle.C16, nseqthresh=0, alenthresh=0)
    assert np.isclose(mi[1], g_ref / (2 * N), rtol=1e-7)


KW0 = dict(nseqthresh=0, alenthresh=0)


def _gapped_two_column(rng, N=300, gap=0.15, other=0.03):
    ax = rng.integers(0, 4, size=(N, 2)).astype(np.uint8)
    ax[:, 1] = np.where(rng.random(N) < 0.6, (ax[:, 0] + 3) % 4, ax[:, 1])     # some covariation
    ax[rng.random((N, 2)) < gap] = 4
    ax[rng.random((N, 2)) < other] = 15
    return ax


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_mig_is_mi_minus_gap_fraction(seed):
    """MIg = MI - ngap / nseff (correlators.c:816): MI from scikit-learn's mutual_info_score on the
    sequences canonical in both columns (L = 2: pm is the pair's own marginal), ngap and nseff
    counted directly."""
    from sklearn.metrics import mutual_info_score
    ax = _gapped_two_column(np.random.default_rng(seed))
    both = (ax[:, 0] < 4) & (ax[:, 1] < 4)
    nseff, ngap = int(both.sum()), int((~both).sum())
    assert ngap > 0
    mi_ref = mutual_info_score(ax[both, 0], ax[both, 1])
    pr = oracle.probs(ax)
    mi, _ = oracle.cov(pr, oracle.MI, oracle.C16, **KW0)
    mig, _ = oracle.cov(pr, oracle.MIg, oracle.C16, **KW0)
    assert np.isclose(mi[1], mi_ref, rtol=1e-7, atol=1e-12)
    assert np.isclose(mig[1], mi_ref - ngap / nseff, rtol=1e-7, atol=1e-12)


@pytest.mark.parametrize("seed", [3, 4])
def test_omes_is_chi2_over_16_for_balanced_columns(seed):
    """OMES = sum (obs - exp)^2 / nseff (correlators.c:227-260).  With every residue exactly N/4
    times in each column the expectation is N/16 in every cell, so OMES = Pearson chi^2 / 16
    (scipy's chi2_contingency)."""
    rng = np.random.default_rng(seed)
    N = 400
    ax = np.stack([rng.permutation(np.repeat(np.arange(4), N // 4)),
                   rng.permutation(np.repeat(np.arange(4), N // 4))], axis=1).astype(np.uint8)
    table = np.zeros((4, 4))
    np.add.at(table, (ax[:, 0], ax[:, 1]), 1)
    chi2 = stats.chi2_contingency(table, correction=False)[0]
    pr = oracle.probs(ax)
    omes, _ = oracle.cov(pr, oracle.OMES, oracle.C16, **KW0)
    assert np.isclose(omes[1], chi2 / 16.0, rtol=1e-7)


def test_ccf_hand_three_columns():
    """CCF (correlators.c:1010-1061) on three gap-free columns of 4 sequences worked by hand:
    compositions pm0 = (1/2, 1/2, 0, 0), pm1 = (1, 0, 0, 0), pm2 = (0, 0, 1/2, 1/2), nseff = 4 for
    every pair, so meanp = (2 pm0 + pm1) / 3 = (2/3, 1/3, 0, 0) (pairs (0,1), (0,2), (1,2)),
    |4 pm0 - meanp| = sqrt(41)/3, |4 pm1 - meanp| = sqrt(101)/3, |4 pm2 - meanp| = sqrt(77)/3,
    and CCF(i, j) = |4 pm_i - meanp| |4 pm_j - meanp|."""
    ax = np.array([[0, 0, 2], [0, 0, 3], [1, 0, 2], [1, 0, 3]], dtype=np.uint8)
    pr = oracle.probs(ax)
    ccf, _ = oracle.cov(pr, oracle.CCF, oracle.C16, **KW0)
    ccf = ccf.reshape(3, 3)
    assert np.isclose(ccf[0, 1], np.sqrt(41 * 101) / 9, rtol=1e-8)
    assert np.isclose(ccf[0, 2], np.sqrt(41 * 77) / 9, rtol=1e-8)
    assert np.isclose(ccf[1, 2], np.sqrt(101 * 77) / 9, rtol=1e-8)
    assert ccf[1, 0] == ccf[0, 1]


@pytest.mark.parametrize("seed", [5, 6, 7])
def test_c2_pooled_tests_match_scipy(seed):
    """C2 (correlators.c:155-173, 463-482, 583-614): cells pooled into allowed (WC + GU) and the
    rest; GT and CHI are then scipy's G-test and Pearson test of the two pooled observed counts
    against the pooled expectations (power_divergence), MI = G / (2 N)."""
    from rscape_amd import ALLOW_WCGU
    rng = np.random.default_rng(seed)
    table = rng.integers(1, 40, size=(4, 4)).astype(float)
    ax = two_column(table.astype(int))
    N = table.sum()
    allow = ALLOW_WCGU.reshape(4, 4) > 0
    q = np.outer(table.sum(1), table.sum(0)) / N ** 2
    obs = np.array([table[allow].sum(), table[~allow].sum()])
    exp = N * np.array([q[allow].sum(), q[~allow].sum()])
    g = stats.power_divergence(obs, exp, lambda_="log-likelihood")[0]
    x2 = stats.power_divergence(obs, exp, lambda_="pearson")[0]
    pr = oracle.probs(ax)
    gt, _ = oracle.cov(pr, oracle.GT, oracle.C2, **KW0)
    chi, _ = oracle.cov(pr, oracle.CHI, oracle.C2, **KW0)
    mi, _ = oracle.cov(pr, oracle.MI, oracle.C2, **KW0)
    assert np.isclose(gt[1], g, rtol=1e-7)
    assert np.isclose(chi[1], x2, rtol=1e-7)
    assert np.isclose(mi[1], g / (2 * N), rtol=1e-7)


@pytest.mark.parametrize("seed", [8, 9])
def test_cwc_partitions_the_g_test(seed):
    """CWC sums the G-test terms of the allowed cells only (correlators.c:397-434): with every
    cell allowed it is scipy's G-test, with none it is 0, and the CWC scores of a cell set and of
    its complement add up to the G-test."""
    from rscape_amd import ALLOW_WCGU
    rng = np.random.default_rng(seed)
    table = rng.integers(1, 40, size=(4, 4))
    ax = two_column(table)
    g = stats.chi2_contingency(table, correction=False, lambda_="log-likelihood")[0]
    pr = oracle.probs(ax)
    cw = lambda ap: oracle.cov(pr, oracle.GT, oracle.CWC, allowpair=ap, **KW0)[0][1]
    assert np.isclose(cw(np.ones(16)), g, rtol=1e-7)
    assert abs(cw(np.zeros(16))) < 1e-12
    assert np.isclose(cw(ALLOW_WCGU) + cw(1.0 - ALLOW_WCGU), g, rtol=1e-7)


def test_bijection_known_answers():
    N = 400
    ax = two_column(np.eye(4) * (N // 4))
    pr = oracle.probs(ax)
    kw = dict(nseqthresh=0, alenthresh=0)
    mi, _ = oracle.cov(pr, oracle.MI, oracle.C16, **kw)
    mir, _ = oracle.cov(pr, oracle.MIr, oracle.C16, **kw)
    gt, _ = oracle.cov(pr, oracle.GT, oracle.C16, **kw)
    chi, _ = oracle.cov(pr, oracle.CHI, oracle.C16, **kw)
    assert np.isclose(mi[1], np.log(4), rtol=1e-8)
    assert np.isclose(mir[1], 1.0, rtol=1e-8)
    assert np.isclose(gt[1], 2 * N * np.log(4), rtol=1e-8)
    assert np.isclose(chi[1], 3 * N, rtol=1e-8)
    ind = two_column(np.ones((4, 4)) * 5)                          # exact independence
    pr = oracle.probs(ind)
    for t in (oracle.MI, oracle.GT, oracle.CHI, oracle.OMES):
        v, _ = oracle.cov(pr, t, oracle.C16, **kw)
        assert abs(v[1]) < 1e-8


def test_apc_asc_hand_matrix():
    L = 4
    M = np.array([[0, 1, 2, 3], [1, 0, 4, 5], [2, 4, 0, 6], [3, 5, 6, 0]], float)
    COV = M.copy()
    np.fill_diagonal(COV, -np.inf)
    X = M.sum(1) / 3.0                           # row means over j != i
    avg = M[np.triu_indices(4, 1)].mean()        # 21 / 6
    apc, mm = oracle.corrected(COV.ravel(), L, oracle.APC)
    asc, _ = oracle.corrected(COV.ravel(), L, oracle.ASC)
    apc, asc = apc.reshape(4, 4), asc.reshape(4, 4)
    for i in range(4):
        for j in range(4):
            if i != j:
                assert np.isclose(apc[i, j], M[i, j] - X[i] * X[j] / avg)
                assert np.isclose(asc[i, j], M[i, j] - (X[i] + X[j] - avg))
    assert np.isneginf(apc[0, 0])
    zero = np.zeros((3, 3)); np.fill_diagonal(zero, -np.inf)
    z, _ = oracle.corrected(zero.ravel(), 3, oracle.APC)                 # COVavg == 0 -> 0
    assert np.all(z.reshape(3, 3)[~np.eye(3, dtype=bool)] == 0)


def test_raf_closed_form_from_counts():
    """RAF's literal O(N^2) loop equals the count closed form the GPU uses."""
    rng = np.random.default_rng(5)
    ax = synth.random_msa(25, 8, rng, gap=0.2)
    N, L = ax.shape
    ref, _ = oracle.raf(ax)
    ref = ref.reshape(L, L)
    from rscape_amd import ALLOW_WCGU
    allow = ALLOW_WCGU.reshape(4, 4) > 0
    c = brute_counts(ax)
    for i in range(L):
        for j in range(i + 1, L):
            t = c[i, j] * allow
            Mv = t.sum()
            cij = 0.5 * (Mv * Mv - (t.sum(1) ** 2).sum()) + 0.5 * (Mv * Mv - (t.sum(0) ** 2).sum())
            cij = cij / (N * (N - 1)) * 2.0
            assert np.isclose(ref[i, j], cij - (N - Mv) / N, rtol=1e-12, atol=1e-14)
    rafs, _ = oracle.raf(ax, rafs=True)
    rafs = rafs.reshape(L, L)
    for i in range(L):
        for j in range(i + 1, L):
            v = 2 * ref[i, j]
            if i > 0 and j < L - 1:
                v += ref[i - 1, j + 1]
            if j > 0 and i < L - 1 and i < j - 2:
                v += ref[i + 1, j - 1]
            assert np.isclose(rafs[i, j], 0.25 * v)


@pytest.mark.parametrize("N,L,seed", [(25, 8, 5), (2, 5, 6), (40, 13, 7), (1, 4, 8)])
def test_raf_counts_matches_literal_loop(N, L, seed):
    """The oracle's count-based RAF (used where the literal loop cannot finish) equals the literal
    O(N^2) loop, and its RAFS stencil equals ref_rafs."""
    rng = np.random.default_rng(seed)
    ax = synth.random_msa(N, L, rng, gap=0.2, other=0.05)
    lit, mm = oracle.raf(ax)
    got, gm = oracle.raf_counts(oracle.pair_counts(ax), N, L)
    off = ~np.eye(L, dtype=bool)
    assert np.allclose(got.reshape(L, L)[off], lit.reshape(L, L)[off], rtol=1e-12, atol=1e-14)
    assert np.allclose(gm, mm, rtol=1e-12, atol=1e-14)
    lrafs, _ = oracle.raf(ax, rafs=True)
    grafs, _ = oracle.rafs_from(got, L)
    assert np.allclose(grafs.reshape(L, L)[off], lrafs.reshape(L, L)[off], rtol=1e-12, atol=1e-14)


def avgid(ax):
    """Mean pairwise identity (percent) over all sequence pairs, identity = identical canonical
    residues / the smaller number of canonical residues of the two sequences."""
    c = ax < 4
    n = ax.shape[0]
    tot = 0.0
    for a in range(n):
        same = (c[a] & c[a + 1:] & (ax[a] == ax[a + 1:])).sum(1)
        ln = np.minimum(c[a].sum(), c[a + 1:].sum(1))
        tot += np.where(ln > 0, same / np.maximum(ln, 1), 0.0).sum()
    return 100.0 * tot / (n * (n - 1) / 2)


def test_tutorial_gtp_scores():
    """Published GTp scores of tutorial/updated_Arisong.sto (documentation/tutorial.tex:202-212)."""
    names, ax, _ = stockholm.read(os.path.join(HERE, "golden", "updated_Arisong.sto"))
    gold = json.load(open(os.path.join(HERE, "golden", "arisong_gtp.json")))
    assert ax.shape == (gold["nseq"], gold["alen_in"])
    use = oracle.gapcol_filter(ax, np.ones(len(names)), 0.75)
    assert use.sum() == gold["alen_analysed"]
    cols = np.nonzero(use)[0]
    ax2 = oracle.canonicalize(ax[:, cols])
    w = oracle.weights_gsc(ax2)
    assert np.isclose(w.sum(), len(names))
    pr = oracle.probs(ax2, w)
    cov, _ = oracle.cov(pr, oracle.GT, oracle.CSELECT)
    L = len(cols)
    cov, mm = oracle.corrected(cov, L, oracle.APC)
    cov = cov.reshape(L, L)
    pos = {c + 1: k for k, c in enumerate(cols)}
    errs = [abs(cov[pos[i], pos[j]] - s) / s for i, j, s in gold["pairs"]]
    # relative error: median 1.5e-5, largest 5.6e-4 (pair 120-139); see the weights note below
    assert max(errs) < 1e-3
    assert np.median(errs) < 5e-5
    assert round(mm[1], 2) == gold["score_range_printed"][1]
    # the header's average identity (documentation/tutorial.tex:158,198: "avgid 65.82 (64.97)")
    # pins the pairwise identity the GSC distances use: identities / min(canonical lengths)
    assert round(avgid(ax2), 2) == 65.82 and round(avgid(oracle.canonicalize(ax)), 2) == 64.97
    # unit weights are far off: the weighting is pinned by these scores
    pr1 = oracle.probs(ax2)
    c1, _ = oracle.cov(pr1, oracle.GT, oracle.CSELECT)
    c1, _ = oracle.corrected(c1, L, oracle.APC)
    assert abs(c1.reshape(L, L)[pos[98], pos[106]] - 121.65824) > 5


def test_pb_weights_hand_cases():
    ax = np.array([[0, 1, 2], [0, 1, 2], [3, 2, 1]], dtype=np.uint8)
    w = oracle.weights_pb(ax)
    assert np.isclose(w.sum(), 3)
    assert np.isclose(w[0], w[1]) and w[2] > w[0]
    # Henikoff: column with r=2 types; seqs 0,1 share a type (n=2): 1/(2*2); seq 2: 1/(2*1)
    raw = np.array([3 * 0.25, 3 * 0.25, 3 * 0.5]) / 3
    assert np.allclose(w, raw / raw.sum() * 3)
    same = np.zeros((5, 4), dtype=np.uint8)
    assert np.allclose(oracle.weights_pb(same), 1.0)


def test_gsc_weights_properties():
    rng = np.random.default_rng(3)
    base = synth.random_msa(1, 40, rng, gap=0.0)
    ax = np.repeat(base, 4, axis=0)
    ax[3, :20] = (ax[3, :20] + 1) % 4                              # one divergent sequence
    w = oracle.weights_gsc(ax)
    assert np.isclose(w.sum(), 4)
    # identical sequences share the weight of their zero-length clade: GSC splits a
    # zero-length pair of subtrees in halves, so the three copies get 1/4, 1/4, 1/2 of it
    assert np.isclose(sorted(w[:3])[0], sorted(w[:3])[1]) and np.isclose(sorted(w[:3])[2], 2 * sorted(w[:3])[0])
    assert w[3] > w[:3].max() and w[:3].sum() < 3


# ---- histogram / fits / E-values ----
def hist_from(values, bmin, bmax, w):
    L = oracle.lib()
    h = L.ref_hist_create(bmin, bmax, w)
    for v in values:
        L.ref_hist_add(C.c_void_p(h), C.c_double(v))
    return h


class Hist(C.Structure):
    _fields_ = [("bmin", C.c_double), ("bmax", C.c_double), ("w", C.c_double), ("nb", C.c_int),
                ("obs", C.POINTER(C.c_uint64)), ("imin", C.c_int), ("imax", C.c_int), ("cmin", C.c_int),
                ("xmin", C.c_double), ("xmax", C.c_double), ("phi", C.c_double),
                ("n", C.c_uint64), ("Nc", C.c_uint64), ("No", C.c_uint64), ("tailfit", C.c_int)]


def H(h):
    return C.cast(h, C.POINTER(Hist)).contents


def test_histogram_binning_is_left_open_right_closed():
    b = C.c_int()
    L = oracle.lib()
    h0 = hist_from([], 0.0, 1.0, 0.05)
    for x, want in [(0.05, 0), (0.0500001, 1), (0.1, 1), (0.0, -1)]:
        L.ref_hist_score2bin(C.c_void_p(h0), C.c_double(x), C.byref(b))
        assert b.value == want, (x, b.value)
    h = hist_from([0.0, 0.05, 0.0500001, 0.1, 2.5, -1.0], 0.0, 1.0, 0.05)
    s = H(h)
    assert s.n == 6 and s.xmax == 2.5 and s.xmin == -1.0
    assert s.bmin < -1.0 and s.bmax >= 2.5                         # grew both ways


def test_exponential_binned_mle_and_tail():
    rng = np.random.default_rng(11)
    x = rng.exponential(1 / 1.7, size=20000) + 0.3
    w = 0.05
    h = hist_from(x, 0.0, x.max() + 5 * w, w)
    L = oracle.lib()
    newmass = C.c_double()
    L.ref_hist_set_tail_by_mass(C.c_void_p(h), C.c_double(0.2), C.byref(newmass))
    s = H(h)
    assert s.tailfit == 1 and 0.2 <= newmass.value < 0.25           # first bin edge past the mass
    mu, lam = C.c_double(), C.c_double()
    L.ref_exp_fit_binned(C.c_void_p(h), C.byref(mu), C.byref(lam))
    assert mu.value == s.phi
    # independent: maximise the binned exponential likelihood numerically
    obs = np.array([s.obs[b] for b in range(s.cmin, s.imax + 1)], float)
    a = s.bmin + w * np.arange(s.cmin, s.imax + 1) - s.phi

    def nll(l):
        return -np.sum(obs * np.log(np.exp(-l * a) - np.exp(-l * (a + w))))
    best = optimize.minimize_scalar(nll, bounds=(0.01, 20), method="bounded", options={"xatol": 1e-10})
    assert np.isclose(lam.value, best.x, rtol=1e-6)
    assert abs(lam.value - 1.7) < 0.1


def test_gamma_fit_matches_scipy_on_midpoints():
    """Gamma fit: PARITY UNPINNED for Easel's binned method; checks the restated ML method."""
    rng = np.random.default_rng(12)
    x = rng.gamma(2.5, 1 / 1.3, size=30000)
    w = 0.02
    h = hist_from(x, 0.0, x.max() + 5 * w, w)
    L = oracle.lib()
    nm = C.c_double()
    L.ref_hist_set_tail_by_mass(C.c_void_p(h), C.c_double(1.0), C.byref(nm))
    s = H(h)
    mu, lam, tau = C.c_double(), C.c_double(), C.c_double()
    L.ref_gam_fit_binned(C.c_void_p(h), C.byref(mu), C.byref(lam), C.byref(tau))
    mids, cnt = [], []
    for b in range(s.cmin, s.imax + 1):
        if s.obs[b]:
            mids.append(s.bmin + w * (b + 0.5) - s.phi)
            cnt.append(s.obs[b])
    data = np.repeat(mids, cnt)
    a, loc, scale = stats.gamma.fit(data, floc=0)
    assert np.isclose(tau.value, a, rtol=1e-4)
    assert np.isclose(lam.value, 1 / scale, rtol=1e-4)
    L.ref_gam_surv.argtypes = [C.c_double] * 4
    for xx in (0.5, 2.0, 5.0):
        assert np.isclose(L.ref_gam_surv(xx, 0.0, 1.3, 2.5), stats.gamma.sf(xx, 2.5, scale=1 / 1.3), rtol=1e-9)


def test_cov2evalue_empirical_and_fit_paths():
    L = oracle.lib()
    vals = [0.1 * k for k in range(100)]
    h = hist_from(vals, -10.0, 15.0, 0.1)
    s = H(h)
    L.ref_cov2evalue.argtypes = [C.c_double, C.c_int64, C.c_void_p, C.c_void_p]
    L.ref_evalue2cov.argtypes = [C.c_double, C.c_int64, C.c_void_p, C.c_void_p]
    # empirical: fraction of the sample at or above the bin of cov
    e = L.ref_cov2evalue(5.05, 1, h, None)
    b = int(np.ceil((5.05 + 10.0) / 0.1 - 1))
    c = sum(s.obs[i] for i in range(b, s.imax + 1))
    assert np.isclose(e, c / s.Nc)
    assert np.isclose(L.ref_cov2evalue(100.0, 7, h, None), 7 / s.Nc)
    # fit path uses survfit[icov + 1]
    nm = C.c_double()
    L.ref_hist_set_tail_by_mass(C.c_void_p(h), C.c_double(0.3), C.byref(nm))
    s = H(h)
    L.ref_survfit.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_void_p]
    ep = np.array([s.phi, 1.0, 0.0])
    sf = L.ref_survfit(h, nm.value, 0, P(ep))
    sfa = np.ctypeslib.as_array(C.cast(sf, C.POINTER(C.c_double)), shape=(2 * s.nb,))
    cov = s.phi + 0.33
    ic = int(np.ceil((cov + 10.0) / 0.1 - 1))
    assert np.isclose(L.ref_cov2evalue(cov, 1, h, C.c_void_p(sf)), sfa[ic + 1])
    ub = s.bmin + 0.1 * (ic + 2)
    assert np.isclose(sfa[ic + 1], nm.value * np.exp(-(ub - s.phi)))
    t = L.ref_evalue2cov(0.01, 1, h, C.c_void_p(sf))
    assert nm.value * np.exp(-(t + 0.1 - s.phi)) <= 0.0105


def test_null_merge_preserves_counts():
    L = oracle.lib()
    L.ref_null_add.argtypes = [C.POINTER(C.c_void_p), C.c_void_p]
    cum = C.c_void_p(None)
    rng = np.random.default_rng(2)
    total = np.zeros(0)
    for r in range(3):
        v = rng.normal(0, 1 + r, 500)
        h = hist_from(np.maximum(v, -10 + 0.05), -10.0, v.max() + 0.25, 0.05)
        L.ref_null_add(C.byref(cum), C.c_void_p(h))
        total = np.concatenate([total, np.maximum(v, -10 + 0.05)])
    s = H(cum.value)
    assert s.Nc == 1500 and s.n == 1500
    obs = np.array([s.obs[b] for b in range(s.nb)])
    ref = np.zeros(s.nb)
    for v in total:
        ref[int(np.ceil((v - s.bmin) / s.w - 1))] += 1
    assert np.array_equal(obs, ref)


def test_width_and_nshuffle_defaults():
    L = oracle.lib()
    assert np.isclose(L.ref_width(30.0, -40.0, -10.0, 0.05, 400, 1e-6), 0.05)
    assert np.isclose(L.ref_width(4.0, 1.0, -10.0, 0.05, 400, 1e-6), 3.0 / 400)
    assert L.ref_width(1e-7, 0.0, -10.0, 0.05, 400, 1e-6) == 0.0
    assert L.ref_nshuffle_default(95, 66) == 20
    assert L.ref_nshuffle_default(10, 60) == 200
    assert L.ref_nshuffle_default(100, 45) == 200
    assert L.ref_nshuffle_default(100, 30) == 300
    assert L.ref_nshuffle_default(3, 300) == 300


def test_philox_known_answers():
    """Random123 Philox4x32-10 known-answer vectors."""
    L = oracle.lib()
    out = (C.c_uint32 * 4)()
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in cases:
        L.ref_philox4x32((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
        assert tuple(out) == want


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    dims = [5, 7, 6, 2]
    nl = 3
    Ws = [rng.normal(0, 0.5, (dims[l + 1], dims[l])) for l in range(nl)]
    bs = [rng.normal(0, 0.1, dims[l + 1]) for l in range(nl)]
    x = rng.normal(size=(4, 5))
    t = rng.normal(size=(4, 2))
    L = oracle.lib()
    d = (C.c_int * 4)(*dims)
    aw = sum(dims[l + 1] * (2 if l < nl - 1 else 1) for l in range(nl))

    def fwd(Ws, bs):
        Wp = (C.c_void_p * nl)(*[P(w) for w in Ws])
        bp = (C.c_void_p * nl)(*[P(b) for b in bs])
        y = np.zeros((4, 2)); acts = np.zeros((4, aw))
        L.ref_mlp_forward(nl, d, Wp, bp, P(x), C.c_int64(4), P(y), P(acts))
        return y, acts, Wp, bp

    y, acts, Wp, bp = fwd(Ws, bs)
    dy = 2 * (y - t)                                    # d/dy of sum (y - t)^2
    dW = [np.zeros_like(w) for w in Ws]; db = [np.zeros_like(b) for b in bs]
    L.ref_mlp_backward(nl, d, Wp, bp, P(x), C.c_int64(4), P(acts), P(dy),
                       (C.c_void_p * nl)(*[P(a) for a in dW]), (C.c_void_p * nl)(*[P(a) for a in db]))
    eps = 1e-6
    for l in range(nl):
        for idx in [(0, 0), (dims[l + 1] - 1, dims[l] - 1)]:
            Wc = [w.copy() for w in Ws]; Wc[l][idx] += eps
            yp, _, _, _ = fwd(Wc, bs)
            Wc[l][idx] -= 2 * eps
            ym, _, _, _ = fwd(Wc, bs)
            num = (np.sum((yp - t) ** 2) - np.sum((ym - t) ** 2)) / (2 * eps)
            assert np.isclose(dW[l][idx], num, rtol=1e-5, atol=1e-8)
        bc = [b.copy() for b in bs]; bc[l][0] += eps
        yp, _, _, _ = fwd(Ws, bc)
        bc[l][0] -= 2 * eps
        ym, _, _, _ = fwd(Ws, bc)
        assert np.isclose(db[l][0], (np.sum((yp - t) ** 2) - np.sum((ym - t) ** 2)) / (2 * eps), rtol=1e-5, atol=1e-8)


def _branches(tree, N):
    for v in range(N - 1):
        for d in (tree["left"][v], tree["right"][v]):
            yield N + v, (-d if d <= 0 else N + d)


def _fitch_score(tree, N, col):
    """Textbook Fitch parsimony score of one column (leaf states col[0..N-1], no gaps)."""
    sets = {}
    score = 0
    for v in range(N - 2, -1, -1):
        kids = []
        for d in (tree["left"][v], tree["right"][v]):
            kids.append({int(col[-d])} if d <= 0 else sets[d])
        inter = kids[0] & kids[1]
        if inter:
            sets[v] = inter
        else:
            sets[v] = kids[0] | kids[1]
            score += 1
    return score


def test_fitch_null_invariants():
    """R-scape's null (Fitch ancestors, one column shuffle, per-branch substitution replay):
    the root's column multiset and every branch's (old -> new) substitution counts survive the
    shuffle, input leaves are kept, and on gap-free columns the reconstruction is most
    parsimonious (its change count equals the Fitch score; brute force on 4 leaves)."""
    from rscape_amd import synth
    rng = np.random.default_rng(12)
    N, L = 40, 90
    tree = synth.yule_tree(N, rng, 0.2)
    ax = synth.evolve(tree, L, rng, (0.25,) * 4)[1].copy()
    ax[rng.random((N, L)) < 0.08] = 4
    ax[rng.random((N, L)) < 0.01] = 15                    # unknown residue: a random state
    allr, sh = oracle.fitch_null_rows(ax, tree, 99)
    assert np.array_equal(allr[:N], ax) and allr[N:].max() <= 4 and sh.max() <= 4
    assert np.array_equal(np.sort(allr[N]), np.sort(sh[N]))
    for pa, ch in _branches(tree, N):
        def subs(a, b):
            m = (a != b) & (a <= 4) & (b <= 4)
            return np.bincount(a[m].astype(int) * 5 + b[m], minlength=25)
        assert np.array_equal(subs(allr[pa], allr[ch]), subs(sh[pa], sh[ch])), (pa, ch)
    for c in range(L):
        if np.all(ax[:, c] < 4):
            changes = sum(int(allr[pa, c] != allr[ch, c]) for pa, ch in _branches(tree, N))
            assert changes == _fitch_score(tree, N, ax[:, c])
    # brute force: 4 leaves, every internal assignment
    import itertools
    t4 = synth.yule_tree(4, rng, 0.2)
    for trial in range(30):
        col = rng.integers(0, 4, size=4).astype(np.uint8)
        a4, _ = oracle.fitch_null_rows(col[:, None], t4, trial)
        got = sum(int(a4[pa, 0] != a4[ch, 0]) for pa, ch in _branches(t4, 4))
        best = 10
        for states in itertools.product(range(4), repeat=3):
            full = list(col) + list(states)
            best = min(best, sum(int(full[pa] != full[ch]) for pa, ch in _branches(t4, 4)))
        assert got == best
