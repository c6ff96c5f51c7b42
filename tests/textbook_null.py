"""Textbook restatement of the simulated null, for distributional checks (test infrastructure).

The generative model of cov_GenerateAlignment's ungapped evolution (src/cov_simulate.c:288-451,
585-631, 724-798): the root drawn from the stationary distribution, each child drawn from row
`parent` of P(t) = expm(t Q) of its branch.  This version takes P(t) from scipy's matrix
exponential and its uniforms from numpy's generator, so it shares neither the Philox keys nor
the eigendecomposed inverse-CDF tables with the device simulator or the oracle.

With one fixed tree and no gaps the columns (for the structured null: the paired column
couples) evolve independently, and a column's marginal pm depends on that column alone, so the
GT scores of disjoint column pairs (0,1), (2,3), ... of one alignment are independent draws of
the null score distribution: a two-sample KS test over them is exact.
"""
import numpy as np
from scipy.linalg import expm

import oracle
from rscape_amd import synth


def evolve_gtr(tree, L, pi, rng, scale=1.0):
    """Leaves (N x L) of ungapped GTR evolution; branch lengths multiplied by scale."""
    n = tree["n"]
    Q = synth.gtr_rate(pi)
    seq = {0: rng.choice(4, size=L, p=np.asarray(pi, float))}
    leaves = np.zeros((n, L), np.uint8)
    for v in range(n - 1):                          # parents come before children
        for child, t in ((tree["left"][v], tree["ld"][v]), (tree["right"][v], tree["rd"][v])):
            cdf = np.cumsum(np.clip(expm(Q * t * scale), 0, None), 1)
            cdf /= cdf[:, -1:]
            c = np.minimum((rng.random(L)[:, None] > cdf[seq[v]]).sum(1), 3)
            if child > 0:
                seq[int(child)] = c
            else:
                leaves[-int(child)] = c
    return leaves


def evolve_pairs16(tree, npairs, Q16, M16, rng, scale=1.0):
    """Leaves (N x 2 npairs) where columns (2k, 2k+1) evolve as one 16-state character
    (state a * 4 + b): root from M16, children from expm(t Q16)."""
    n = tree["n"]
    seq = {0: rng.choice(16, size=npairs, p=np.asarray(M16, float) / np.sum(M16))}
    leaves = np.zeros((n, 2 * npairs), np.uint8)
    for v in range(n - 1):
        for child, t in ((tree["left"][v], tree["ld"][v]), (tree["right"][v], tree["rd"][v])):
            cdf = np.cumsum(np.clip(expm(Q16 * t * scale), 0, None), 1)
            cdf /= cdf[:, -1:]
            c = np.minimum((rng.random(npairs)[:, None] > cdf[seq[v]]).sum(1), 15)
            if child > 0:
                seq[int(child)] = c
            else:
                leaves[-int(child), 0::2] = c // 4
                leaves[-int(child), 1::2] = c % 4
    return leaves


def disjoint_pair_gt(cov, L):
    """GT of the column pairs (2k, 2k+1) from a full L x L COV matrix."""
    C = np.asarray(cov).reshape(L, L)
    k = np.arange(L // 2)
    return C[2 * k, 2 * k + 1]


def oracle_gt(ax):
    """GT (16-cell class) of every column pair by the oracle."""
    L = ax.shape[1]
    COV, _ = oracle.cov(oracle.probs(ax), oracle.GT, oracle.C16)
    return disjoint_pair_gt(COV, L)


def pairing(L):
    """ct with columns (2k, 2k+1) paired."""
    ct = np.arange(L, dtype=np.int32) ^ 1
    return ct
