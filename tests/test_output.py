"""Output files of a significance run (host code of the library, no GPU): the sorted hit list and
the power block reproduce what documentation/tutorial.tex:193-235 publishes for
tutorial/updated_Arisong.sto when given the published per-pair values; the power table reader
follows power_Read; the oracle's substitution counts are pinned by Fitch parsimony scores."""
import json
import os

import numpy as np
import pytest

import oracle
from rscape_amd import output as O
from rscape_amd import synth

HERE = os.path.join(os.path.dirname(__file__), "golden")
GOLD = json.load(open(os.path.join(HERE, "arisong_output.json")))
ALEN = 150                                     # positions are reported in original coordinates
WSUM = 95.0                                    # GSC weights sum to nseq


def _nseff(printed, pct):
    """An nseff within the rounding interval of both printed values (nseff to 2 decimals, its
    share of the 95 weight units to 1 decimal): the published values are known to that precision."""
    lo = max(printed - 0.005, (pct - 0.05) * WSUM / 100.0)
    hi = min(printed + 0.005, (pct + 0.05) * WSUM / 100.0)
    assert lo < hi, (printed, pct)
    return 0.5 * (lo + hi)


def _spairs(rows):
    """SPair array over all 150 columns with the given (i, j, nsubs, power, nseff, canonical%,
    covary) rows filled (1-based positions); ngap = 95 - nseff gives the published nseff(%)."""
    dim = ALEN * (ALEN - 1) // 2
    arr = (O.SPair * dim)()
    for n in range(dim):
        arr[n].powertype = O.SINGLE_SUBS
        arr[n].bptype_given = O.BP_NONE
    for i, j, ns, pw, nseff, can, cov in rows:
        a, b = i - 1, j - 1
        n = a * ALEN - a * (a + 1) // 2 + (b - a - 1)
        s = arr[n]
        s.i, s.j, s.iabs, s.jabs = a, b, i, j
        s.nsubs, s.power, s.nseff, s.ngap = ns, pw, nseff, WSUM - nseff
        s.wcf_f, s.wobble_f = can / 100.0, 0.0
        s.bptype_given = O.BP_WWC
        s.covary = int(cov)
    return arr


def test_sorted_hit_list_reproduces_tutorial(tmp_path):
    rows = []
    hits = []
    for t in GOLD["hits"]:
        i, j, sc, ev, pv, ns, pw, can = int(t[1]), int(t[2]), float(t[3]), float(t[4]), float(t[5]), int(t[6]), \
            float(t[7]), float(t[10])
        nseff = _nseff(float(t[8]), float(t[9].strip("(%)")))
        rows.append((i, j, ns, pw, nseff, can, True))
        h = O.Hit()
        h.i, h.j, h.sc, h.Eval, h.pval, h.nsubs, h.power, h.bptype, h.is_compatible = i - 1, j - 1, sc, ev, pv, ns, pw, \
            O.BP_WWC, 0
        hits.append(h)
    spair = _spairs(rows)
    # the hits in pair order (as cov_CreateHitList makes them); the ranked file sorts by E-value
    hits.sort(key=lambda h: (h.i, h.j))
    f = str(tmp_path / "x.sorted.cov")
    O.write_hits(f, hits, spair, ALEN, "GTp", 0.05, -9.95, 121.66, tnum=20, Nb=20, Nt=0, ranked=True,
                 samplesize=O.SAMPLE_BP)
    lines = open(f).read().splitlines()
    summ = next(l for l in lines if l.startswith("# GTp"))
    assert summ.split() == GOLD["summary"]
    got = [l.split() for l in lines if l.startswith("*")]
    assert got == GOLD["hits"]
    # every data line keeps the reference's tab-separated layout
    assert all(l.count("\t") == 10 for l in lines if l.startswith("*"))


def test_power_block_reproduces_tutorial(tmp_path):
    block = GOLD["power_block"]
    rows = []
    for t in block:
        if t[0] == "*" and len(t) == 8:
            rows.append((int(t[1]), int(t[2]), int(t[3]), float(t[4]), _nseff(float(t[5]), float(t[6].strip("(%)"))),
                         float(t[7]), True))
        elif t[0].isdigit():
            rows.append((int(t[0]), int(t[1]), int(t[2]), float(t[3]), _nseff(float(t[4]), float(t[5].strip("(%)"))),
                         float(t[6]), False))
    assert len(rows) == 20
    spair = _spairs(rows)
    f = str(tmp_path / "x.power")
    O.write_power(f, spair, ALEN * (ALEN - 1) // 2, True)
    got = [l.split() for l in open(f).read().splitlines() if l.strip()]
    assert got[0] == ["#", "Power", "analysis", "of", "given", "structure"]
    assert got[2:] == block                   # every line, including the expected-covariation summary


def test_power_table_reader(tmp_path):
    """power_Read: powers clamped to [0, 1] and held at 1 after the first row above 1; the power of
    a pair is the last row's whose substitution count it exceeds."""
    f = tmp_path / "p.csv"
    f.write_text("0 0.0\n2 -0.1\n5 0.25\n9 0.5\n14 1.2\n20 0.9\n")
    p = O.PowerTable(str(f))
    subs, prob = p.table()
    assert list(subs) == [0, 2, 5, 9, 14, 20] and list(prob) == [0.0, 0.0, 0.25, 0.5, 1.0, 1.0]
    assert [p.prob(n) for n in (0, 1, 3, 6, 10, 15, 21, 100)] == [0.0, 0.0, 0.0, 0.25, 0.5, 1.0, 1.0, 1.0]


def _fitch_score(ax, tree, col):
    """Textbook Fitch parsimony score of one gap-free column (set intersections and unions)."""
    N = ax.shape[0]
    sets, score = {}, 0

    def node(d):
        return sets[d] if d > 0 else {int(ax[-d, col])}
    for v in range(N - 2, -1, -1):
        a, b = node(tree["left"][v]), node(tree["right"][v])
        if a & b:
            sets[v] = a & b
        else:
            sets[v] = a | b
            score += 1
    return score


@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_substitutions_are_parsimony_scores(seed):
    """Without gaps every column's substitution count is its Fitch parsimony score, whatever the
    tie-breaks, and njoin = nsubs_i + nsubs_j - ndouble (inclusion-exclusion)."""
    rng = np.random.default_rng(seed)
    N, L = 40, 25
    tree = synth.yule_tree(N, rng, 0.3)
    ax = synth.evolve(tree, L, rng)[1]
    ns, nd, nj = oracle.tree_substitutions(ax, tree, seed=5 + seed)
    assert [int(v) for v in ns] == [_fitch_score(ax, tree, c) for c in range(L)]
    iu = np.triu_indices(L, 1)
    assert np.array_equal(nj[iu], (ns[:, None] + ns[None, :])[iu] - nd[iu])
    assert np.all(nd[iu] <= np.minimum(ns[:, None], ns[None, :])[iu])
    ns2, nd2, nj2 = oracle.tree_substitutions(ax, tree, seed=5 + seed, includegaps=True)
    assert np.array_equal(ns2, ns) and np.array_equal(nd2, nd) and np.array_equal(nj2, nj)


def test_oracle_substitutions_gap_rule():
    """Two leaves "A-" and "CA": column 0 changes on exactly one branch; at column 1 the root is A
    (gaps do not enter the union) and the branch to the gapped leaf counts only with gaps."""
    tree = dict(left=np.array([0], np.int32), right=np.array([-1], np.int32), ld=np.ones(1), rd=np.ones(1), n=2)
    ax = np.array([[0, 4], [1, 0]], np.uint8)
    for seed in range(4):
        ns, _, _ = oracle.tree_substitutions(ax, tree, seed=seed)
        assert list(ns) == [1, 0]
        ns, _, _ = oracle.tree_substitutions(ax, tree, seed=seed, includegaps=True)
        assert list(ns) == [1, 1]
