"""GPU null simulator and batched null pipeline.

Simulation is checked symbol for symbol against the oracle (same Philox keys), P(t) tables
against scipy's matrix exponential, and distributional properties (stationary composition,
substitutions per site).  The batched pipeline is checked against the single-alignment
path (which is itself checked against the oracle), eager against hipGraph replays.
"""
import numpy as np
import pytest
from scipy.linalg import expm

import oracle
import rscape_amd as R
from rscape_amd import null as NL

pytestmark = pytest.mark.gpu
PI = (0.25, 0.22, 0.30, 0.23)


def test_gtr_cdf_matches_expm(ctx):
    from rscape_amd import synth
    m = NL.gtr_model(PI)
    ts = np.array([0.0, 0.01, 0.1, 0.5, 2.0, 10.0])
    got = NL.gtr_cdf(ctx, m, ts).astype(np.int64)
    Q = synth.gtr_rate(PI)
    for k, t in enumerate(ts):
        P = expm(Q * t)
        assert np.allclose(P.sum(1), 1.0)
        ref = np.floor(np.cumsum(P, 1)[:, :3] * 2.0 ** 32)
        assert np.max(np.abs(got[k] - np.minimum(ref, 2 ** 32 - 1))) <= 4, t


@pytest.mark.parametrize("N,L,gap", [(2, 5, 0.0), (37, 64, 0.0), (300, 130, 0.1)])
def test_simulation_matches_oracle(ctx, N, L, gap):
    m = NL.gtr_model(PI)
    tree = NL.yule_tree(N, 0.15, seed=N)
    job = NL.Job(tree, L, m, seed=1234 + N, stream=7, gapfrac=gap)
    got = NL.simulate(ctx, job)
    gapthr = 0 if gap <= 0 else int(gap * 2 ** 32)
    ref = oracle.evolve_gtr(tree, L, m, 1234 + N, 7, gapthr)
    assert np.mean(got != ref) < 1e-5
    assert got.max() <= 4


def test_simulation_distribution(ctx):
    """Leaf composition is stationary; mean pairwise difference follows the tree."""
    m = NL.gtr_model(PI)
    tree = NL.yule_tree(400, 0.1, seed=3)
    leaves = NL.simulate(ctx, NL.Job(tree, 2000, m, seed=9))
    comp = np.bincount(leaves.ravel(), minlength=4)[:4] / leaves.size
    assert np.allclose(comp, PI, atol=0.01)
    # a 2-leaf tree with branch t on each side: P(differ) = 1 - sum_a pi_a [P(2t)]_aa
    from rscape_amd import synth
    t = 0.3
    two = dict(left=np.array([0], np.int32), right=np.array([-1], np.int32), ld=np.array([t]), rd=np.array([t]), n=2)
    lv = NL.simulate(ctx, NL.Job(two, 200000, m, seed=5))
    P = expm(synth.gtr_rate(PI) * 2 * t)
    expect = 1 - float(np.sum(np.array(PI) * np.diag(P)))
    assert abs(np.mean(lv[0] != lv[1]) - expect) < 0.005


@pytest.mark.parametrize("kind", ["star", "rand"])
def test_star_and_random_trees(ctx, kind):
    """R-scape-sim's star and random trees: leaves exchangeable, each pair of leaves the same as
    on a two-leaf tree, P(same symbol) = sum_a pi_a [P(2t)]_aa for a star of leaf branch t and
    sum_a pi_a^2 for the random tree (independent residues); composition stationary."""
    from rscape_amd import synth
    m = NL.gtr_model(PI)
    N, L, t = 300, 3000, 0.3
    tree = NL.star_tree(N, t) if kind == "star" else NL.rand_tree(N)
    inner = np.concatenate([tree["ld"][tree["left"] > 0], tree["rd"][tree["right"] > 0]])
    assert np.all(inner == 0) and len(inner) == N - 2                # the star's zero-length spine
    lv = NL.simulate(ctx, NL.Job(tree, L, m, seed=21))
    assert lv.shape == (N, L) and lv.max() <= 3
    comp = np.bincount(lv.ravel(), minlength=4)[:4] / lv.size
    assert np.allclose(comp, PI, atol=0.01)
    P = expm(synth.gtr_rate(PI) * 2 * t) if kind == "star" else np.tile(np.array(PI), (4, 1))
    expect = float(np.sum(np.array(PI) * np.diag(P)))
    one = np.eye(4, dtype=np.float32)[lv].reshape(N, 4 * L)   # one-hot rows
    same = (one @ one.T).astype(np.float64) / L              # P(same) for every leaf pair
    iu = np.triu_indices(N, 1)
    sd = np.sqrt(expect * (1 - expect) / L)
    assert abs(same[iu].mean() - expect) < 0.003
    # exchangeable: no pair further from the expectation than chance allows (about 45k pairs)
    assert np.max(np.abs(same[iu] - expect)) < 6 * sd
    # two leaves under one node are no closer than two from opposite halves of the tree
    assert abs(same[0, 1] - same[0, N - 1]) < 6 * sd


def single_path_fit(ctx, leaves, covtype="GTp", pmass=5e-4, expo=True):
    msa = ctx.msa(leaves)
    mn, mx = msa.calculate(covtype)
    w = R.width(mx, mn)
    h = msa.rank_hist(-10.0, w)
    fit, sf = h.fit(ctx, pmass, 1.0, expo)
    return fit, w, mn, mx, h


def make_jobs(sizes, seed=0):
    m = NL.gtr_model(PI)
    jobs = []
    for k, (N, L) in enumerate(sizes):
        tree = NL.yule_tree(N, 0.05 + 0.05 * (k % 5), seed=seed + k)
        jobs.append(NL.Job(tree, L, m, seed=seed * 1000 + k, stream=k, gapfrac=0.05 * (k % 2)))
    return jobs


def test_batch_matches_single_path(ctx):
    sizes = [(50, 80), (300, 211), (9, 64), (120, 47), (1000, 96)]
    jobs = make_jobs(sizes, seed=11)
    fits, hs = NL.batch(ctx, jobs, "GTp", pmass=0.05, hists=True)
    for k, job in enumerate(jobs):
        leaves = NL.simulate(ctx, job)
        fit, w, mn, mx, h = single_path_fit(ctx, leaves, pmass=0.05)
        print("job", k, sizes[k], "batch", fits[k][:6], "single", fit, w)
        assert np.isclose(fits[k][5], w, rtol=1e-12)
        assert np.isclose(fits[k][6], mn, rtol=1e-9, atol=1e-9) and np.isclose(fits[k][7], mx, rtol=1e-9)
        assert np.array_equal(hs[k].h.counts(), h.h.counts()[:hs[k].h.nb])
        for i, key in enumerate(["mu", "lambda", "tau", "newmass", "pmass"]):
            assert np.isclose(fits[k][i], fit[key], rtol=1e-9, atol=1e-12), key


def test_batch_persistent_equals_per_tile(ctx):
    """The persistent queue Gram and the one-group-per-tile Gram give identical nulls, with
    large (64 x 32-column) and small (32 x 32) tiles in one batch."""
    print("persistent vs per-tile batch", flush=True)
    sizes = [(300, 300), (1000, 420), (50, 80), (700, 257), (9, 64), (2000, 90)]
    jobs = make_jobs(sizes, seed=21)
    ctx.set_batch_persistent(False)
    try:
        a, ha = NL.batch(ctx, jobs, "GTp", pmass=0.05, hists=True)
    finally:
        ctx.set_batch_persistent(True)
    b, hb = NL.batch(ctx, jobs, "GTp", pmass=0.05, hists=True)
    assert np.array_equal(np.nan_to_num(a), np.nan_to_num(b))
    for x, y in zip(ha, hb):
        assert np.array_equal(x.h.counts(), y.h.counts())


def test_batch_graph_replays_equal_eager(ctx):
    sizes = [(64, 100), (200, 150), (33, 80), (500, 60)]
    jobs = make_jobs(sizes, seed=5)
    eager, he = NL.batch(ctx, jobs, "GTp", graph=0, hists=True)
    for reps in (1, 3):
        g, hg = NL.batch(ctx, jobs, "GTp", graph=reps, hists=True)
        _ = ctx.msa(np.zeros((4, 4), np.uint8))            # eager work between calls
        assert np.array_equal(np.nan_to_num(g), np.nan_to_num(eager))
        for a, b in zip(he, hg):
            assert np.array_equal(a.h.counts(), b.h.counts())


def test_batch_graph_cache_replays_new_inputs(ctx):
    """A graph call of a geometry already captured replays the cached graph with the new
    call's inputs (seeds, branch lengths, models): equal to an eager run of those inputs.  A
    call of another geometry captures afresh."""
    print("graph cache: simulated batch", flush=True)
    sizes = [(64, 100), (200, 150), (33, 80), (500, 60)]
    first, _ = NL.batch(ctx, make_jobs(sizes, seed=5), "GTp", graph=1)
    g0 = ctx.graph_reuses()
    jobs = make_jobs(sizes, seed=5)
    # same geometry (the tree shapes repeat), other branch lengths, models and seeds
    for k, j in enumerate(jobs):
        j.tree["ld"] = j.tree["ld"] * 1.7
        j.tree["rd"] = j.tree["rd"] * 0.6
        j.model = np.ascontiguousarray(NL.gtr_model((0.4, 0.1, 0.2, 0.3)))
        j.seed += 99
    eager, he = NL.batch(ctx, jobs, "GTp", graph=0, hists=True)
    g, hg = NL.batch(ctx, jobs, "GTp", graph=1, hists=True)
    assert ctx.graph_reuses() == g0 + 1
    assert not np.array_equal(np.nan_to_num(g), np.nan_to_num(first))
    assert np.array_equal(np.nan_to_num(g), np.nan_to_num(eager))
    for a, b in zip(he, hg):
        assert np.array_equal(a.h.counts(), b.h.counts())
    other, _ = NL.batch(ctx, make_jobs(sizes[:3], seed=6), "GTp", graph=1)
    assert ctx.graph_reuses() == g0 + 1
    eager3, _ = NL.batch(ctx, make_jobs(sizes[:3], seed=6), "GTp", graph=0)
    assert np.array_equal(np.nan_to_num(other), np.nan_to_num(eager3))
    print("graph cache: R-scape null batch", flush=True)
    from rscape_amd import synth
    rng = np.random.default_rng(3)
    tree = synth.yule_tree(80, rng, 0.15)
    ax = synth.evolve(tree, 120, rng, (0.3, 0.2, 0.2, 0.3))[1].copy()
    NL.batch_fitch(ctx, ax, tree, [1, 2, 3], "GTp", pmass=0.01, graph=1)
    r0 = ctx.graph_reuses()
    ax2 = synth.evolve(tree, 120, rng, (0.25, 0.25, 0.25, 0.25))[1].copy()
    fit_e, h_e = NL.batch_fitch(ctx, ax2, tree, [7, 8, 9], "GTp", pmass=0.01)
    fit_g, h_g = NL.batch_fitch(ctx, ax2, tree, [7, 8, 9], "GTp", pmass=0.01, graph=1)
    assert ctx.graph_reuses() == r0 + 1
    assert np.array_equal(fit_g, fit_e) and np.array_equal(h_g.h.counts(), h_e.h.counts())


def test_batch_groups_match_cumulative_null(ctx):
    """A group of replicates = R-scape's null loop: width from the first replicate, every
    replicate's histogram merged (null_add2cumranklist), one fit on the merged histogram."""
    m1, m2 = NL.gtr_model(PI), NL.gtr_model((0.4, 0.1, 0.1, 0.4))
    tree_a, tree_b = NL.yule_tree(120, 0.1, seed=1), NL.yule_tree(40, 0.25, seed=2)
    jobs = [NL.Job(tree_a, 90, m1, seed=77, stream=r, gapfrac=0.05, group=0) for r in range(5)]
    jobs += [NL.Job(tree_b, 150, m2, seed=78, stream=r, group=1) for r in range(3)]
    jobs += [NL.Job(tree_a, 61, m2, seed=79, stream=0, group=2)]
    fits, hs = NL.batch(ctx, jobs, "GTp", pmass=0.01, hists=True)
    assert fits.shape == (3, 8)
    for g, members in enumerate([jobs[:5], jobs[5:8], jobs[8:]]):
        cum, w = None, None
        mn, mx = np.inf, -np.inf
        for job in members:
            msa = ctx.msa(NL.simulate(ctx, job))
            a, b = msa.calculate("GTp")
            mn, mx = min(mn, a), max(mx, b)
            if w is None:
                w = R.width(b, a)
            h = msa.rank_hist(-10.0, w)
            cum = R.Histogram.null_add(cum, h)
        fit, _ = cum.fit(ctx, 0.01, 1.0, True)
        assert fits[g][5] == w and fits[g][6] == mn and fits[g][7] == mx
        got = hs[g].h
        assert (got.n, got.Nc, got.imin, got.imax) == (cum.h.n, cum.h.Nc, cum.h.imin, cum.h.imax)
        assert np.array_equal(got.counts()[got.imin:got.imax + 1], cum.h.counts()[cum.h.imin:cum.h.imax + 1])
        for i, key in enumerate(["mu", "lambda", "tau", "newmass", "pmass"]):
            assert np.isclose(fits[g][i], fit[key], rtol=1e-9, atol=1e-12), (g, key)


def test_batch_rejects_bad_groups_and_trees(ctx):
    m = NL.gtr_model(PI)
    t = NL.yule_tree(10, 0.1)
    with pytest.raises(R.RscapeError):
        NL.batch(ctx, [NL.Job(t, 20, m, group=1)])
    with pytest.raises(R.RscapeError):
        NL.batch(ctx, [NL.Job(t, 20, m, group=0), NL.Job(t, 20, m, group=2)])
    c = t["left"][0] if t["left"][0] > 0 else t["right"][0]
    bad = dict(t); bad["left"] = t["left"].copy(); bad["right"] = t["right"].copy()
    bad["left"][0] = bad["right"][0] = c                                   # one internal node reached twice
    with pytest.raises(R.RscapeError):
        NL.simulate(ctx, NL.Job(bad, 20, m))


@pytest.mark.parametrize("N,L,seed", [(2, 7, 1), (4, 40, 2), (60, 130, 3), (500, 300, 4), (24, 4100, 5)])
def test_fitch_null_matches_oracle(ctx, N, L, seed):
    """R-scape's null on the device: symbol-identical to the oracle on the same keys."""
    from rscape_amd import synth
    rng = np.random.default_rng(seed)
    tree = synth.yule_tree(N, rng, 0.15)
    ax = synth.evolve(tree, L, rng, (0.3, 0.2, 0.2, 0.3))[1].copy()
    ax[rng.random((N, L)) < 0.1] = 4
    ax[rng.random((N, L)) < 0.02] = 15
    got = NL.fitch_null(ctx, ax, tree, seed=77 + seed)
    _, sh = oracle.fitch_null_rows(ax, tree, 77 + seed)
    assert np.array_equal(got, sh[:N])


@pytest.mark.parametrize("N,L,nrep", [(60, 130, 4), (300, 211, 3), (5, 80, 1), (9, 64, 2)])
def test_batch_fitch_matches_cumulative_null(ctx, N, L, nrep):
    """R-scape's own null loop in one batch (every Fitch kernel covers all replicates) = the
    per-replicate path: rsg_null_fitch per seed, score, width from the first replicate,
    histograms merged, one fit; eager and graph replays agree."""
    from rscape_amd import synth
    print("batch fitch N=%d L=%d nrep=%d" % (N, L, nrep), flush=True)
    rng = np.random.default_rng(N + L)
    tree = synth.yule_tree(N, rng, 0.15)
    ax = synth.evolve(tree, L, rng, (0.3, 0.2, 0.2, 0.3))[1].copy()
    ax[rng.random((N, L)) < 0.05] = 4
    ax[rng.random((N, L)) < 0.01] = 15
    seeds = [1000 + 17 * r for r in range(nrep)]
    fit_b, h_b = NL.batch_fitch(ctx, ax, tree, seeds, "GTp", pmass=0.01)
    cum, w = None, None
    mn, mx = np.inf, -np.inf
    for sd in seeds:
        msa = ctx.msa(NL.fitch_null(ctx, ax, tree, seed=sd))
        a, b = msa.calculate("GTp")
        mn, mx = min(mn, a), max(mx, b)
        if w is None:
            w = R.width(b, a)
        cum = R.Histogram.null_add(cum, msa.rank_hist(-10.0, w))
    fit, _ = cum.fit(ctx, 0.01, 1.0, True)
    assert fit_b[5] == w and fit_b[6] == mn and fit_b[7] == mx
    got = h_b.h
    assert (got.n, got.Nc, got.imin, got.imax) == (cum.h.n, cum.h.Nc, cum.h.imin, cum.h.imax)
    assert np.array_equal(got.counts()[got.imin:got.imax + 1], cum.h.counts()[cum.h.imin:cum.h.imax + 1])
    for i, key in enumerate(["mu", "lambda", "tau", "newmass", "pmass"]):
        assert np.isclose(fit_b[i], fit[key], rtol=1e-9, atol=1e-12), key
    fit_g, h_g = NL.batch_fitch(ctx, ax, tree, seeds, "GTp", pmass=0.01, graph=2)
    assert np.array_equal(fit_g, fit_b)
    assert np.array_equal(h_g.h.counts(), h_b.h.counts())


def _pairing(L, npairs, seed):
    rng = np.random.default_rng(seed)
    perm = rng.permutation(L)[:2 * npairs]
    ct = -np.ones(L, np.int32)
    for k in range(npairs):
        i, j = int(perm[2 * k]), int(perm[2 * k + 1])
        ct[i], ct[j] = j, i
    return ct


@pytest.mark.parametrize("N,L,npairs,gap", [(2, 9, 3, 0.0), (37, 64, 20, 0.0), (300, 131, 40, 0.1)])
def test_pair_simulation_matches_oracle(ctx, N, L, npairs, gap):
    """Structured simulation (16-state pairs): symbol-identical to the oracle on the same keys."""
    from test_pair_model import pair_params
    print("pair sim N=%d L=%d pairs=%d" % (N, L, npairs), flush=True)
    m = NL.gtr_model(PI)
    Q, M = pair_params(N)
    m16 = NL.pair_model(Q, M)
    tree = NL.yule_tree(N, 0.15, seed=N)
    ct = _pairing(L, npairs, N)
    job = NL.Job(tree, L, m, seed=99 + N, stream=3, gapfrac=gap)
    got = NL.simulate_pairs(ctx, job, ct, m16)
    gapthr = 0 if gap <= 0 else int(gap * 2 ** 32)
    ref = oracle.evolve_pairs(tree, L, m, m16, ct, 99 + N, 3, gapthr)
    assert np.mean(got != ref) < 1e-5
    assert got.max() <= 4


def test_pair_simulation_distribution(ctx):
    """Long branches: leaf pair states of paired columns follow the stationary pair frequencies
    (so base pairs covary: WC/GU fraction = its weight in M16), unpaired columns follow pi."""
    from test_pair_model import pair_params, WC
    m = NL.gtr_model(PI)
    Q, M = pair_params(7, wc=0.85)
    m16 = NL.pair_model(Q, M)
    tree = NL.yule_tree(300, 1.0, seed=4)
    L, npairs = 3000, 1000
    ct = _pairing(L, npairs, 5)
    leaves = NL.simulate_pairs(ctx, NL.Job(tree, L, m, seed=21), ct, m16)
    i5 = np.array([i for i in range(L) if ct[i] > i]); i3 = ct[i5]
    st = 4 * leaves[:, i5].astype(int) + leaves[:, i3]
    freq = np.bincount(st.ravel(), minlength=16) / st.size
    assert np.allclose(freq, M, atol=0.01)
    wc = sum(M[a * 4 + b] for a, b in WC)
    assert abs(np.isin(st, [a * 4 + b for a, b in WC]).mean() - wc) < 0.01
    un = leaves[:, ct < 0]
    assert np.allclose(np.bincount(un.ravel(), minlength=4)[:4] / un.size, PI, atol=0.01)


def _device_gt(ctx, ax):
    msa = ctx.msa(ax)
    try:
        msa.probs(R.COUNTS_INT8)
        msa.cov("GT", R.C16)
        return msa.get_cov()[0]
    finally:
        msa.close()


@pytest.mark.parametrize("structured", [False, True])
def test_simulated_null_scores_ks(ctx, structured):
    """Device-simulated nulls validated by distribution: GT scores of the disjoint column pairs
    (0,1), (2,3), ... of 100 device-simulated alignments on one tree (independent draws: columns
    evolve independently without gaps), scored on the device, against the textbook null
    (scipy expm, numpy's generator) scored by the oracle -- two-sample KS; the same test
    rejects a textbook null with every branch 1.25x too long."""
    from scipy.stats import ks_2samp
    from rscape_amd import synth
    from test_pair_model import pair_params
    import textbook_null as TB
    N, L, Rn = 40, 64, 100
    print("KS null structured=%d" % structured, flush=True)
    tree = synth.yule_tree(N, np.random.default_rng(8), mean_bl=0.1)
    m = NL.gtr_model(PI)
    Q, M = pair_params(4, wc=0.8)
    m16, ct = NL.pair_model(Q, M), TB.pairing(L)
    got = []
    for r in range(Rn):
        job = NL.Job(tree, L, m, seed=505, stream=r)
        ax = NL.simulate_pairs(ctx, job, ct, m16) if structured else NL.simulate(ctx, job)
        got.append(TB.disjoint_pair_gt(_device_gt(ctx, ax), L))
    got = np.concatenate(got)

    def textbook(seed, scale=1.0):
        rng = np.random.default_rng(seed)
        if structured:
            return np.concatenate([TB.oracle_gt(TB.evolve_pairs16(tree, L // 2, Q, M, rng, scale)) for _ in range(Rn)])
        return np.concatenate([TB.oracle_gt(TB.evolve_gtr(tree, L, PI, rng, scale)) for _ in range(Rn)])

    assert ks_2samp(got, textbook(41)).pvalue > 1e-3
    assert ks_2samp(got, textbook(42, scale=1.25)).pvalue < 1e-3


def _rates8(r):
    rI, muAM, ldI, muI = r
    return [rI, 0.0, 0.0, muAM, muAM, muAM, ldI, muI]


@pytest.mark.parametrize("N,L,rates,bl,use_root,npairs,gap", [
    (2, 7, NL.AIF_PFAM, 0.1, False, 0, 0.0),
    (5, 40, (0.5, 0.5, 0.3, 0.6), 0.5, True, 0, 0.0),
    (40, 120, NL.AIF_PFAM, 0.15, True, 0, 0.0),
    (40, 120, NL.AIF_PFAM, 0.15, False, 12, 0.05),
    (200, 120, NL.AIF_PFAM, 0.05, True, 30, 0.0),
    (300, 500, NL.AIF_PFAM, 0.05, False, 0, 0.0),
])
def test_generate_alignment_matches_oracle(ctx, N, L, rates, bl, use_root, npairs, gap):
    """cov_GenerateAlignment with the e1 indel model on the device = the oracle's sequential
    restatement of the reference loop (ungapped evolution from the given or drawn root, then the
    branch-by-branch indel walks growing the alignment) on the same keys: same length, same
    base-pair map, same gap pattern, residues identical up to the rare P(t)-table rounding the
    ungapped simulators allow.  Every branch walks the whole alignment of its moment, inserted
    columns included, so the length grows geometrically with the number of branches; the
    branch lengths keep it to a few times the root's."""
    from test_pair_model import pair_params
    print("generate alignment N=%d L=%d rates=%s root=%s pairs=%d" % (N, L, rates, use_root, npairs), flush=True)
    rng = np.random.default_rng(N + L)
    m = NL.gtr_model(PI)
    tree = NL.yule_tree(N, bl, seed=N + 1)
    job = NL.Job(tree, L, m, seed=555 + N, stream=2, gapfrac=gap)
    root = rng.integers(0, 4, size=L, dtype=np.uint8) if use_root else None
    gapthr = 0 if gap <= 0 else int(gap * 2 ** 32)
    if npairs:
        Q, M = pair_params(N)
        m16 = NL.pair_model(Q, M)
        ct = _pairing(L, npairs, N)
        got, gct = NL.generate_alignment(ctx, job, root=root, ct=ct, model16=m16, rates=NL.aif_rates(*rates))
        ung = oracle.evolve_pairs(tree, L, m, m16, ct, 555 + N, 2, gapthr, root=root)
    else:
        ct = None
        got, gct = NL.generate_alignment(ctx, job, root=root, rates=NL.aif_rates(*rates))
        ung = oracle.evolve_gtr(tree, L, m, 555 + N, 2, gapthr, root=root)
    ref, rct = oracle.add_indels(tree, ung, _rates8(rates), m[36:40], 555 + N, 2, ct=ct)
    print("alignment length %d (root %d)" % (ref.shape[1], L), flush=True)
    assert got.shape == ref.shape and np.array_equal(gct, rct)
    assert np.array_equal(got == 4, ref == 4)
    assert np.mean(got != ref) < 1e-5
    # without indels: the ungapped leaves from the root
    plain, pct = NL.generate_alignment(ctx, job, root=root, ct=ct, model16=m16 if npairs else None)
    assert plain.shape == (N, L) and np.mean(plain != ung) < 1e-5
    if use_root:
        assert np.array_equal(pct, -np.ones(L, np.int32) if ct is None else ct)


def test_generate_alignment_rejects_bad_input(ctx):
    m = NL.gtr_model(PI)
    tree = NL.yule_tree(6, 0.1, seed=1)
    job = NL.Job(tree, 10, m, seed=1)
    with pytest.raises(R.RscapeError, match="not canonical"):
        NL.generate_alignment(ctx, job, root=np.full(10, 4, np.uint8))
    with pytest.raises(R.RscapeError):
        NL.generate_alignment(ctx, job, rates=NL.aif_rates(1.0, 0.1, 0.1, 0.3))
