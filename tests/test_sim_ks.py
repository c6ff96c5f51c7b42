"""The oracle's simulated null checked by distribution (CPU).

Two-sample KS tests on GT null scores: the oracle's simulator (Philox keys, eigendecomposed
inverse-CDF tables; the checker of the device simulator) against the textbook restatement in
textbook_null.py (scipy expm, numpy's generator).  The same tests must reject a textbook null
with every branch 1.25x too long, so a mis-scaled P(t) or a wrong branch cannot pass.
"""
import numpy as np
from scipy.stats import ks_2samp

import oracle
import textbook_null as TB
from rscape_amd import null as NL
from rscape_amd import synth

PI = (0.25, 0.22, 0.30, 0.23)
N, L, R = 40, 64, 100          # 32 independent pair scores per replicate, 3200 per sample
P_ACCEPT = 1e-3


def _tree(seed=5):
    return synth.yule_tree(N, np.random.default_rng(seed), mean_bl=0.1)


def test_gtr_null_scores_match_textbook():
    tree, m = _tree(), NL.gtr_model(PI)
    a = np.concatenate([TB.oracle_gt(oracle.evolve_gtr(tree, L, m, 77, r)) for r in range(R)])
    rng = np.random.default_rng(11)
    b = np.concatenate([TB.oracle_gt(TB.evolve_gtr(tree, L, PI, rng)) for _ in range(R)])
    assert ks_2samp(a, b).pvalue > P_ACCEPT
    rng = np.random.default_rng(12)
    bad = np.concatenate([TB.oracle_gt(TB.evolve_gtr(tree, L, PI, rng, scale=1.25)) for _ in range(R)])
    assert ks_2samp(a, bad).pvalue < P_ACCEPT


def test_pair_null_scores_match_textbook():
    from test_pair_model import pair_params
    tree, m = _tree(6), NL.gtr_model(PI)
    Q, M = pair_params(4, wc=0.8)
    m16 = NL.pair_model(Q, M)
    ct = TB.pairing(L)
    a = np.concatenate([TB.oracle_gt(oracle.evolve_pairs(tree, L, m, m16, ct, 31, r)) for r in range(R)])
    rng = np.random.default_rng(21)
    b = np.concatenate([TB.oracle_gt(TB.evolve_pairs16(tree, L // 2, Q, M, rng)) for _ in range(R)])
    assert ks_2samp(a, b).pvalue > P_ACCEPT
    rng = np.random.default_rng(22)
    bad = np.concatenate([TB.oracle_gt(TB.evolve_pairs16(tree, L // 2, Q, M, rng, scale=1.25)) for _ in range(R)])
    assert ks_2samp(a, bad).pvalue < P_ACCEPT
