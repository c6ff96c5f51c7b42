"""GPU: Tree_Substitutions on the device against the oracle (exact integers), and at the
N=5000 x L=1500 size against a linear-algebra form of the counts; the per-pair summary of
power_SPAIR_Create against the tutorial's published canonical-pair percentages; the hit list
against the oracle's E-values; the .surv file."""
import json
import os

import numpy as np
import pytest

import oracle
import rscape_amd as R
from rscape_amd import output as O
from rscape_amd import stockholm, synth

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("N,L,gap,incl", [(2, 5, 0.0, 0), (40, 33, 0.1, 0), (40, 33, 0.1, 1), (300, 130, 0.15, 0),
                                          (97, 300, 0.05, 1)])
def test_substitutions_match_oracle(ctx, N, L, gap, incl):
    print("substitutions N=%d L=%d gap=%.2f includegaps=%d" % (N, L, gap, incl), flush=True)
    rng = np.random.default_rng(N + L)
    tree = synth.yule_tree(N, rng, 0.2)
    ax = synth.evolve(tree, L, rng, (0.3, 0.2, 0.2, 0.3))[1].copy()
    ax[rng.random((N, L)) < gap] = 4
    ax[rng.random((N, L)) < 0.02] = 15
    ns, nd, nj = O.substitutions(ctx, ax, tree, seed=11, includegaps=incl)
    rns, rnd, rnj = oracle.tree_substitutions(ax, tree, seed=11, includegaps=incl)
    assert np.array_equal(ns, rns)
    assert np.array_equal(nd, rnd) and np.array_equal(nj, rnj)


def test_substitutions_full_size(ctx):
    """N=5000 x L=1500 (9998 branches): nsubs exact against the oracle; ndouble = D^T D and
    njoin = C^T C - E^T E over the oracle's Fitch states, with D = substitution, E = no
    substitution, C = D + E (parent and child canonical) per branch and column."""
    d = synth.config(1, seed=42)
    ax, tree = d["ax"], d["tree"]
    N, L = ax.shape
    ns, nd, nj = O.substitutions(ctx, ax, tree, seed=3)
    allr, _ = oracle.fitch_null_rows(ax, tree, 3)
    left, right = np.asarray(tree["left"]), np.asarray(tree["right"])
    rows = lambda c: np.where(c > 0, N + c, -c)
    pa = np.repeat(np.arange(N - 1) + N, 2)
    ch = np.stack([rows(left), rows(right)], 1).reshape(-1)
    P, Cc = allr[pa], allr[ch]
    can = (P < 4) & (Cc < 4)
    D = (can & (P != Cc)).astype(np.float32)
    E = (can & (P == Cc)).astype(np.float32)
    assert np.array_equal(ns, D.sum(0).astype(np.int32))
    iu = np.triu_indices(L, 1)
    assert np.array_equal(nd[iu], (D.T @ D)[iu].astype(np.int32))
    Cm = D + E
    assert np.array_equal(nj[iu], (Cm.T @ Cm - E.T @ E)[iu].astype(np.int32))
    assert np.all(np.tril(nd) == 0) and np.all(np.tril(nj) == 0)


def test_pair_summary_matches_tutorial_canonical(ctx):
    """power_SPAIR_AddMI's canonical-pair share, 100 (wcf + wobble) of pp, on the tutorial
    alignment with device GSC weights: the % canonical_pairs column the tutorial publishes for the
    20 pairs of the structure (documentation/tutorial.tex:215-235), and nseff / ngap against the
    oracle."""
    names, ax, _ = stockholm.read(os.path.join(HERE, "updated_Arisong.sto"))
    gold = json.load(open(os.path.join(HERE, "arisong_output.json")))
    use, _ = R.gapcol_mask(ax, np.ones(len(names)), 0.75)
    cols = np.nonzero(use)[0]
    sub = oracle.canonicalize(np.ascontiguousarray(ax[:, cols]))
    w = ctx.msaweight(sub)
    m = ctx.msa(sub, w)
    m.probs(R.COUNTS_BF16X3)
    Lf = len(cols)
    sp = O.spairs(ctx, m, nsubs=np.zeros(Lf, np.int32), msamap=cols)
    pos = {c + 1: k for k, c in enumerate(cols)}
    n_of = lambda a, b: a * Lf - a * (a + 1) // 2 + (b - a - 1)
    checked = 0
    for t in gold["power_block"]:
        if t[0] == "*" and len(t) == 8:
            i, j, can = int(t[1]), int(t[2]), t[7]
        elif t[0].isdigit():
            i, j, can = int(t[0]), int(t[1]), t[6]
        else:
            continue
        s = sp[n_of(pos[i], pos[j])]
        assert (s.iabs, s.jabs) == (i, j)
        assert "%.1f" % (100.0 * (s.wcf_f + s.wobble_f)) == can, (i, j)
        checked += 1
    assert checked == 20
    pr = oracle.probs(sub, w)
    ns_ref, ng_ref = pr["nseff"].reshape(Lf, Lf), pr["ngap"].reshape(Lf, Lf)
    for n in (0, 17, 500, len(sp) - 1):
        s = sp[n]
        assert np.isclose(s.nseff, ns_ref[s.i, s.j], rtol=1e-5) and np.isclose(s.ngap, ng_ref[s.i, s.j], rtol=1e-4, atol=1e-4)
    m.close()


def test_hit_list_matches_oracle_evalues(ctx, tmp_path):
    """cov_CreateHitList's selection (one-set test: E = pval * #pairs) against the oracle's
    E-values on a fitted null; the .cov and .sorted.cov files list the same pairs; .surv written."""
    from rscape_amd import null as NL
    d = synth.config(0, seed=13)
    ax, tree = d["ax"], d["tree"]
    N, L = ax.shape
    Nc = L * (L - 1) // 2
    fit, h = NL.batch_fitch(ctx, ax, tree, [5, 6, 7, 8], "GTp", pmass=0.05)
    _, sf = h.fit(ctx, 0.05, 1.0, True)
    m = ctx.msa(ax)
    m.calculate("GTp")
    cov = m.get_cov()[0]
    sp = O.spairs(ctx, m, nsubs=np.zeros(L, np.int32))
    hits = O.hit_list(ctx, m, h, sf, O.SAMPLE_ALL, Nt=Nc, thresh=0.05, spair=sp)
    ev = m.evalues(h, sf, Nc)
    iu = np.triu_indices(L, 1)
    ref = sorted((int(i), int(j)) for i, j in zip(*iu) if ev[i, j] < 0.05)
    assert [(h_.i, h_.j) for h_ in hits] == ref
    for h_ in hits:
        assert np.isclose(h_.Eval, ev[h_.i, h_.j], rtol=1e-12) and h_.sc == cov[h_.i, h_.j]
    assert sum(s.covary for s in sp) == len(hits)
    f1, f2, f3 = str(tmp_path / "a.cov"), str(tmp_path / "a.sorted.cov"), str(tmp_path / "a.surv")
    O.write_hits(f1, hits, sp, L, "GTp", 0.05, float(h.h.xmin), float(h.h.xmax), Nt=Nc)
    O.write_hits(f2, hits, sp, L, "GTp", 0.05, float(h.h.xmin), float(h.h.xmax), Nt=Nc, ranked=True)
    a = sorted(tuple(l.split()[1:3]) for l in open(f1) if not l.startswith("#"))
    b = sorted(tuple(l.split()[1:3]) for l in open(f2) if not l.startswith("#"))
    assert a == b == sorted((str(i + 1), str(j + 1)) for i, j in ref)
    O.write_surv(f3, null=h, survfit=sf)
    blocks = open(f3).read().split("&\n")
    obs = [tuple(map(float, l.split())) for l in blocks[0].splitlines()]
    assert obs and all(obs[k][0] >= obs[k + 1][0] for k in range(len(obs) - 1))
    assert abs(obs[-1][1] - 1.0) < 1e-9                    # the survival reaches every null pair
    m.close()
