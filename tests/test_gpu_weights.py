"""Sequence weights on the device against the oracle.

GSC: the identity counts are integers and the UPGMA/GSC arithmetic is the same FP64 sequence,
so the weights match the oracle to the last bit, including alignments with many tied
distances (duplicated sequences).  PB: the per-sequence sums are reduced in another order,
so they agree to 1e-12 relative.  The tutorial alignment pins GSC end to end.
"""
import json
import os

import numpy as np
import pytest

import oracle
from rscape_amd import stockholm

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def rand_msa(rng, N, L, gap=0.1, degen=0.0):
    ax = rng.integers(0, 4, size=(N, L), dtype=np.uint8)
    ax[rng.random((N, L)) < gap] = 4
    if degen:
        ax[rng.random((N, L)) < degen] = rng.integers(5, 18, size=1)[0]
    return ax


@pytest.mark.parametrize("N,L,seed", [(1, 10, 0), (2, 5, 1), (63, 40, 2), (64, 33, 3), (65, 100, 4), (300, 257, 5)])
def test_gsc_matches_oracle(ctx, N, L, seed):
    rng = np.random.default_rng(seed)
    ax = rand_msa(rng, N, L)
    got = ctx.weights(ax, "GSC")
    ref = oracle.weights_gsc(ax)
    assert np.array_equal(got, ref), np.max(np.abs(got - ref))
    assert abs(got.sum() - N) < 1e-9 * N


def test_gsc_ties_and_duplicates(ctx):
    rng = np.random.default_rng(7)
    base = rand_msa(rng, 40, 60, gap=0.05)
    ax = np.concatenate([base, base[:15], base[5:25]], axis=0)   # exact duplicates -> D = 0 ties
    ax = ax[rng.permutation(len(ax))]
    got = ctx.weights(ax, "GSC")
    assert np.array_equal(got, oracle.weights_gsc(ax))
    ax2 = np.zeros((50, 30), np.uint8)                           # all identical: equal weights
    assert np.allclose(ctx.weights(ax2, "GSC"), 1.0)


@pytest.mark.parametrize("N,L,seed", [(1, 10, 0), (7, 3, 1), (500, 300, 2), (5000, 1500, 3)])
def test_pb_matches_oracle(ctx, N, L, seed):
    rng = np.random.default_rng(seed)
    ax = rand_msa(rng, N, L, gap=0.2, degen=0.02)
    got = ctx.weights(ax, "PB")
    ref = oracle.weights_pb(ax)
    assert np.allclose(got, ref, rtol=1e-12, atol=0)


def test_pb_all_gap_rows(ctx):
    ax = np.full((6, 8), 4, np.uint8)
    ax[0, :4] = [0, 1, 2, 3]
    got = ctx.weights(ax, "PB")
    assert np.allclose(got, oracle.weights_pb(ax), rtol=1e-12)


def test_tutorial_scores_through_device_weights(ctx):
    """GSC weights on the device + gap filter + GTp reproduce the tutorial's published scores."""
    import rscape_amd as R
    names, ax, _ = stockholm.read(os.path.join(GOLD, "updated_Arisong.sto"))
    gold = json.load(open(os.path.join(GOLD, "arisong_gtp.json")))
    use, kept = R.gapcol_mask(ax, np.ones(len(names)), 0.75)
    assert kept == gold["alen_analysed"]
    assert np.array_equal(use, oracle.gapcol_filter(ax, np.ones(len(names)), 0.75))
    sub = np.ascontiguousarray(ax[:, use])
    w = ctx.msaweight(sub)
    assert np.array_equal(w, oracle.weights_gsc(sub))
    msa = ctx.msa(sub, w)
    msa.calculate("GTp", counts=R.COUNTS_BF16X3)
    cov, _, _ = msa.get_cov()
    cols = np.nonzero(use)[0]
    pos = {c + 1: k for k, c in enumerate(cols)}
    errs = [abs(cov[pos[i], pos[j]] - s) / s for i, j, s in gold["pairs"]]
    assert max(errs) < 1e-3 and np.median(errs) < 5e-5          # as the oracle (test_oracle.py)


@pytest.mark.parametrize("N,L,seed", [(1, 7, 0), (40, 90, 1), (513, 300, 2)])
def test_device_gap_column_filter(ctx, N, L, seed):
    """msamanip_RemoveGapColumns on the resident alignment: the column decisions equal the
    oracle's (unit and random weights; residues, gaps, degenerate/N, nonresidue and missing
    codes), and the kept-column alignment built on the device scores as one uploaded from the
    host (counts bit-exact, GTp equal)."""
    import rscape_amd as R
    rng = np.random.default_rng(seed)
    ax = rng.integers(0, 4, size=(N, L), dtype=np.uint8)
    ax[rng.random((N, L)) < rng.random(L)[None, :] * 0.8] = 4
    for code, p in ((9, 0.02), (15, 0.02), (16, 0.02), (17, 0.03)):
        ax[rng.random((N, L)) < p] = code
    for w in (None, rng.uniform(0.1, 3.0, size=N)):
        m = ctx.msa(ax, w)
        for thr in (0.0, 0.5, 0.75):
            sub, use, mp = m.remove_gap_columns(thr)
            ref = oracle.gapcol_filter(ax, np.ones(N) if w is None else w, thr)
            assert np.array_equal(use.astype(bool), ref.astype(bool)), (thr, w is None)
            assert np.array_equal(mp, np.nonzero(ref)[0])
            if sub.Lc < 2:
                sub.close()
                continue
            host = ctx.msa(np.ascontiguousarray(ax[:, ref.astype(bool)]), w)
            kw = {} if w is None else dict(counts=R.COUNTS_BF16X3)
            a = sub.calculate("GTp", **kw)
            b = host.calculate("GTp", **kw)
            assert a == b
            assert np.array_equal(sub.counts(), host.counts())
            assert np.array_equal(sub.get_cov()[0], host.get_cov()[0])
            sub.close(); host.close()
        whole, use, mp = m.remove_gap_columns(1.0)            # no filter: every column kept
        assert whole.Lc == L and use.all() and np.array_equal(mp, np.arange(L))
        whole.close(); m.close()
    gaps = ctx.msa(np.full((5, 9), 4, np.uint8))
    with pytest.raises(R.RscapeError, match="no positions left"):
        gaps.remove_gap_columns(0.5)
