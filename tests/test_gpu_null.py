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
