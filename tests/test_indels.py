"""The e1 indel model of cov_GenerateAlignmentAddGaps, on the CPU.

The oracle's transitions are pinned by the textbook closed forms they restate (the TKF91
insertion probability beta(t) and gamma(t) = 1 - exp(-mu t)), the library's host transitions
equal the oracle's, and the oracle's walk is pinned by the moments of the Markov chain: per
column a deletion with probability gamma, per slot between columns a run of insertions of mean
beta / (1 - t_II)."""
import math

import numpy as np
import pytest

import oracle
from rscape_amd import null as NL

AIF = NL.AIF_PFAM          # rI, muAM, ldI, muI


def _rates(rI, muAM, ldI, muI):
    return [rI, 0.0, 0.0, muAM, muAM, muAM, ldI, muI]


def _tkf_beta(ld, mu, t):
    """TKF91 (Thorne, Kishino & Felsenstein 1991): probability that a link has >= 1 descendant
    insertion, lambda (1 - e^{(lambda - mu) t}) / (mu - lambda e^{(lambda - mu) t}); lambda t /
    (1 + lambda t) when lambda = mu."""
    if ld == mu:
        return ld * t / (1 + ld * t)
    e = math.exp((ld - mu) * t)
    return ld * (1 - e) / (mu - ld * e)


@pytest.mark.parametrize("rates,t", [(AIF, 0.1), (AIF, 1.5), ((0.5, 0.5, 0.3, 0.6), 0.5), ((0.2, 0.05, 0.4, 0.4), 0.8),
                                     ((0.3, 0.1, 0.7, 0.2), 0.4), (AIF, 0.0)])
def test_e1_transitions_closed_forms(rates, t):
    rI, muAM, ldI, muI = rates
    o = oracle.e1_transitions(_rates(*rates), t)
    beta, gamma = _tkf_beta(ldI, muI, t), 1 - math.exp(-muAM * t)
    for r in range(4):
        assert abs(sum(o[4 * r:4 * r + 3]) - 1) < 1e-12 and o[4 * r + 3] == 0
    for r in (0, 1, 2):                            # B, S, D (rM = rD = 0)
        assert np.allclose(o[4 * r:4 * r + 3], [(1 - beta) * (1 - gamma), (1 - beta) * gamma, beta], rtol=1e-12, atol=1e-15)
    assert np.allclose(o[12:15], [(1 - rI) * (1 - beta) * (1 - gamma), (1 - rI) * (1 - beta) * gamma, (1 - rI) * beta + rI],
                       rtol=1e-12, atol=1e-15)
    lib = NL.e1_transitions(NL.aif_rates(*rates), t)
    assert np.array_equal(lib, o)


def _star2(t_left):
    """Root with two leaves; only the left branch (walked first) has length."""
    return dict(n=2, left=np.array([0], np.int32), right=np.array([-1], np.int32),
                ld=np.array([t_left]), rd=np.array([0.0]))


def test_indel_walk_moments():
    """Walks over L0 columns of many independent branches: deletions per column and insertions
    per slot match the chain's expectations (within 5 standard errors)."""
    rI, muAM, ldI, muI = 0.5, 0.5, 0.3, 0.6
    t, L0, R = 0.5, 100, 600
    o = oracle.e1_transitions(_rates(rI, muAM, ldI, muI), t)
    # a column is consumed by S or D; the D share of that step is gamma from every state
    gamma, beta, tII = o[5] / (o[4] + o[5]), o[6], o[14]
    assert abs(gamma - (1 - math.exp(-muAM * t))) < 1e-12
    rng = np.random.default_rng(0)
    dels, ins = [], []
    for r in range(R):
        leaves = rng.integers(0, 4, size=(2, L0), dtype=np.uint8)
        out, _ = oracle.add_indels(_star2(t), leaves, _rates(rI, muAM, ldI, muI), (0.25,) * 4, 100 + r, 3)
        orig = out[1] < 4                           # the right leaf is untouched: its residues mark the L0 columns
        assert orig.sum() == L0 and np.array_equal(out[1][orig], leaves[1])
        left = out[0]
        d = int((left[orig] == 4).sum())
        assert np.array_equal(left[orig][left[orig] < 4], leaves[0][left[orig] < 4])
        dels.append(d)
        ins.append(int((left[~orig] < 4).sum()))
        assert (~orig).sum() == ins[-1]             # every inserted column holds the left leaf's residue
    dels, ins = np.array(dels), np.array(ins)
    e_del, e_ins = L0 * gamma, (L0 + 1) * beta / (1 - tII)
    assert abs(dels.mean() - e_del) < 5 * dels.std() / math.sqrt(R), (dels.mean(), e_del)
    assert abs(ins.mean() - e_ins) < 5 * ins.std() / math.sqrt(R), (ins.mean(), e_ins)


def _is_subsequence(a, b):
    it = iter(b)
    return all(any(x == y for y in it) for x in a)


def test_indel_pass_keeps_residues_and_pairs():
    """Original residues are A/C and inserted ones G/U: every output column is either an original
    column (residues A/C only, base pairs kept) or an inserted one (one G/U from a leaf's branch, or
    none from an internal branch's; unpaired), and each leaf's A/C residues are its input row with
    some deleted."""
    rng = np.random.default_rng(5)
    N, L0 = 12, 60
    tree = NL.yule_tree(N, 0.4, seed=3)
    leaves = rng.integers(0, 2, size=(N, L0), dtype=np.uint8)
    ct = -np.ones(L0, np.int32)
    for a, b in [(2, 50), (3, 49), (10, 30), (11, 29), (40, 55)]:
        ct[a], ct[b] = b, a
    out, cto = oracle.add_indels(tree, leaves, _rates(0.4, 0.3, 0.3, 0.5), (0.0, 0.0, 0.5, 0.5), 17, 0, ct=ct)
    F = out.shape[1]
    orig = []
    for c in range(F):
        col = out[:, c]
        res = col[col < 4]
        if len(res) and np.all(res <= 1):
            orig.append(c)
        else:                                       # inserted: by a leaf's branch (one G/U) or an internal one
            assert len(res) <= 1 and (len(res) == 0 or res[0] >= 2) and cto[c] == -1
    assert len(orig) == L0 and F > L0
    orig = np.array(orig)
    paired = np.nonzero(cto >= 0)[0]
    assert np.array_equal(np.sort(paired), np.sort(orig[np.nonzero(ct >= 0)[0]]))
    assert np.array_equal(cto[orig[ct >= 0]], orig[ct[ct >= 0]])
    for l in range(N):
        o = out[l, orig]
        assert np.array_equal(o[o < 4], leaves[l][o < 4])
        assert _is_subsequence(out[l][out[l] <= 1], leaves[l])
