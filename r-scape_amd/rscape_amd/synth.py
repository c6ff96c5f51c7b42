"""Synthetic RNA alignments for parity tests and the benchmark.

The generator follows the recipe R-scape-sim uses (src/R-scape-sim.c:608-680,
src/cov_simulate.c:60-190): a random ultrametric tree rescaled to a target mean
branch length, a root sequence evolved down the tree with P(t) = exp(tQ), paired
columns drawn together from a pair-state chain, then gaps added to the leaves.

Trees use the Easel ESL_TREE convention: N leaves, internal nodes 0..N-2 with root 0
and parents numbered before children; left[v] / right[v] > 0 name internal nodes,
<= 0 name leaf -child; ld[v] / rd[v] are the branch lengths.
"""
import numpy as np
from scipy.linalg import expm

SYMBOLS = "ACGU-RYMKSWHBVDN*~"


def yule_tree(n, rng, mean_bl=0.1):
    """Random Yule (pure-birth) ultrametric tree with n leaves, rescaled so the mean
    branch length is mean_bl (esl_tree_er_RescaleAverageBL semantics)."""
    if n < 2:
        raise ValueError("a tree needs at least 2 leaves")
    # grow from the root: each step splits a uniformly chosen lineage
    left = np.zeros(n - 1, dtype=np.int32)
    right = np.zeros(n - 1, dtype=np.int32)
    birth = np.zeros(n - 1)            # time each internal node was created
    lineages = [("root", 0)]           # open lineages (parent node, side)
    t = 0.0
    parent_of_new = []
    nxt = 0
    # process: k lineages, waiting time ~ Exp(k)
    lin = [(-1, 0)]                    # (parent internal node, side 0/1); root lineage
    node_time = {}
    while nxt < n - 1:
        k = len(lin)
        t += rng.exponential(1.0 / k)
        idx = rng.integers(k)
        par, side = lin.pop(idx)
        v = nxt
        nxt += 1
        birth[v] = t
        if par >= 0:
            if side == 0:
                left[par] = v
            else:
                right[par] = v
        lin.append((v, 0))
        lin.append((v, 1))
    t_end = t + rng.exponential(1.0 / len(lin))
    leaf = 0
    for (par, side) in lin:
        if side == 0:
            left[par] = -leaf
        else:
            right[par] = -leaf
        leaf += 1
    ld = np.zeros(n - 1)
    rd = np.zeros(n - 1)
    for v in range(n - 1):
        for arr, child in ((ld, left[v]), (rd, right[v])):
            tc = birth[child] if child > 0 else t_end
            arr[v] = tc - birth[v]
    scale = mean_bl / np.mean(np.concatenate([ld, rd]))
    return dict(left=left, right=right, ld=ld * scale, rd=rd * scale, n=n)


def gtr_rate(pi, exch=(1.0, 2.0, 1.0, 1.0, 2.0, 1.0)):
    """4-state GTR rate matrix normalised to one expected substitution per unit time.
    exch: exchangeabilities (AC, AG, AU, CG, CU, GU)."""
    pi = np.asarray(pi, dtype=float)
    S = np.zeros((4, 4))
    pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    for (a, b), e in zip(pairs, exch):
        S[a, b] = S[b, a] = e
    Q = S * pi[None, :]
    np.fill_diagonal(Q, -Q.sum(1))
    Q /= -(pi * np.diag(Q)).sum()
    return Q


PAIR_STATES = [(0, 3), (3, 0), (1, 2), (2, 1), (2, 3), (3, 2)]  # AU UA CG GC GU UG


def pair_rate(pi):
    """6-state rate matrix over the WC/GU pair states: single-nucleotide moves at rate 1,
    compensatory double changes at rate 0.5; stationary distribution from pi_x pi_y."""
    pi = np.asarray(pi, dtype=float)
    st = PAIR_STATES
    w = np.array([pi[a] * pi[b] for a, b in st])
    w /= w.sum()
    Q = np.zeros((6, 6))
    for i, (a, b) in enumerate(st):
        for j, (c, d) in enumerate(st):
            if i == j:
                continue
            nd = (a != c) + (b != d)
            Q[i, j] = (1.0 if nd == 1 else 0.5) * w[j]
    np.fill_diagonal(Q, -Q.sum(1))
    Q /= -(w * np.diag(Q)).sum()
    return Q, w


def plant_pairs(L, npairs, rng, hmin=4, hmax=7, loopmin=3):
    """Nested base pairs in helices of hmin..hmax bp; returns ct (partner or -1)."""
    ct = -np.ones(L, dtype=np.int64)
    placed = 0
    tries = 0
    while placed < npairs and tries < 100000:
        tries += 1
        h = int(rng.integers(hmin, hmax + 1))
        h = min(h, npairs - placed)
        free = ct < 0
        # contiguous free runs
        runs = []
        s = None
        for k in range(L + 1):
            f = k < L and free[k]
            if f and s is None:
                s = k
            elif not f and s is not None:
                runs.append((s, k))
                s = None
        runs = [(a, b) for a, b in runs if b - a >= 2 * h + loopmin]
        if not runs:
            break
        a, b = runs[rng.integers(len(runs))]
        span = int(rng.integers(2 * h + loopmin, b - a + 1))
        st = int(rng.integers(a, b - span + 1))
        en = st + span - 1
        for k in range(h):
            ct[st + k] = en - k
            ct[en - k] = st + k
        placed += h
    return ct


def plant_crossing(ct, npairs, rng, hmin=4, hmax=7, loopmin=3, tries=100000):
    """Pseudoknotted base pairs added to a nested ct: helices of hmin..hmax bp between free
    columns, each crossing at least one pair already in ct (i < k < j < l or k < i < l < j).
    Returns (ct, list of the added (i, j) pairs)."""
    ct = ct.copy()
    L = len(ct)
    added = []
    for _ in range(tries):
        if len(added) >= npairs:
            break
        h = min(int(rng.integers(hmin, hmax + 1)), npairs - len(added))
        a = int(rng.integers(0, L - h))
        b = int(rng.integers(a + 2 * h + loopmin - 1, L)) if a + 2 * h + loopmin - 1 < L else -1
        if b < 0:
            continue
        five = list(range(a, a + h))
        three = list(range(b, b - h, -1))
        if any(ct[k] >= 0 for k in five + three):
            continue
        ks = np.nonzero(ct > np.arange(L))[0]             # 5' ends of the pairs so far
        ls = ct[ks]
        i, j = a, b                                       # the helix's outermost pair
        if not np.any(((i < ks) & (ks < j) & (j < ls)) | ((ks < i) & (i < ls) & (ls < j))):
            continue
        for i, j in zip(five, three):
            ct[i], ct[j] = j, i
            added.append((i, j))
    return ct, added


def evolve(tree, L, rng, pi=(0.25, 0.25, 0.25, 0.25), ct=None, exch=(1.0, 2.0, 1.0, 1.0, 2.0, 1.0)):
    """Evolve a root sequence down the tree; returns (all sequences (2N-1) x L, leaves N x L).
    Rows 0..N-2 are internal nodes, row N-1+l is leaf l (cov_simulate.c:288-324 layout)."""
    n = tree["n"]
    pi = np.asarray(pi, dtype=float)
    Q = gtr_rate(pi, exch)
    Qp, wp = pair_rate(pi)
    if ct is None:
        ct = -np.ones(L, dtype=np.int64)
    five = np.nonzero((ct >= 0) & (ct > np.arange(L)))[0]
    unp = np.nonzero(ct < 0)[0]
    seqs = np.zeros((2 * n - 1, L), dtype=np.uint8)
    pst = np.zeros((2 * n - 1, len(five)), dtype=np.int8)
    seqs[0, unp] = rng.choice(4, size=len(unp), p=pi)
    pst[0] = rng.choice(6, size=len(five), p=wp)
    pa = np.array(PAIR_STATES)
    cache = {}

    def cdf(t):
        key = round(float(t), 12)
        if key not in cache:
            P = np.clip(expm(Q * t), 0, None)
            Pp = np.clip(expm(Qp * t), 0, None)
            cache[key] = (np.cumsum(P / P.sum(1, keepdims=True), 1), np.cumsum(Pp / Pp.sum(1, keepdims=True), 1))
        return cache[key]

    for v in range(n - 1):
        for child, t in ((tree["left"][v], tree["ld"][v]), (tree["right"][v], tree["rd"][v])):
            row = child if child > 0 else n - 1 - child
            C1, C2 = cdf(t)
            u = rng.random(len(unp))
            seqs[row, unp] = (u[:, None] > C1[seqs[v, unp]]).sum(1)
            u = rng.random(len(five))
            pst[row] = (u[:, None] > C2[pst[v]]).sum(1)
    three = ct[five]
    seqs[:, five] = pa[pst, 0]
    seqs[:, three] = pa[pst, 1]
    return seqs, np.ascontiguousarray(seqs[n - 1:])


def add_gaps(ax, frac, rng, mean_run=1.0):
    """Gap a fraction of leaf symbols, i.i.d. (mean_run = 1) or in runs of geometric length."""
    ax = ax.copy()
    N, L = ax.shape
    if frac <= 0:
        return ax
    if mean_run <= 1.0:
        ax[rng.random((N, L)) < frac] = 4
        return ax
    p = 1.0 / mean_run
    nstarts = rng.random((N, L)) < frac / mean_run
    lens = rng.geometric(p, size=(N, L))
    s_idx, c_idx = np.nonzero(nstarts)
    for s, c in zip(s_idx, c_idx):
        ax[s, c:c + lens[s, c]] = 4
    return ax


def config(cfg, seed=42, N=None, L=None):
    """Alignments of the benchmark configurations (sizes can be overridden):
    0: N=200 L=120, 30 pairs, mean branch 0.1, 5% i.i.d. gaps, uniform pi;
    1: N=5000 L=1500, 400 pairs, pi A.25 C.22 G.30 U.23, mean branch 0.15, 10% gaps in runs (mean 4);
    4: N=20000 L=5000, 1200 nested pairs + 40 pseudoknotted (crossing) pairs, mean branch 0.2,
       15% gaps in runs."""
    rng = np.random.default_rng(seed)
    if cfg == 0:
        N, L, npairs, bl, gap, run, pi = N or 200, L or 120, 30, 0.1, 0.05, 1.0, (0.25,) * 4
    elif cfg == 1:
        N, L, npairs, bl, gap, run, pi = N or 5000, L or 1500, 400, 0.15, 0.10, 4.0, (0.25, 0.22, 0.30, 0.23)
    elif cfg == 4:
        N, L, npairs, bl, gap, run, pi = N or 20000, L or 5000, 1200, 0.2, 0.15, 4.0, (0.25, 0.22, 0.30, 0.23)
    else:
        raise ValueError("config %r" % cfg)
    npairs = min(npairs, max(0, (L - 3) // 3))
    tree = yule_tree(N, rng, bl)
    ct = plant_pairs(L, npairs, rng)
    pk = []
    if cfg == 4:                                   # 40 pseudoknotted pairs across the nested ones
        ct, pk = plant_crossing(ct, min(40, max(0, (L - 3) // 20)), rng)
    _, leaves = evolve(tree, L, rng, pi, ct)
    ax = add_gaps(leaves, gap, rng, run)
    return dict(ax=ax, ct=ct, tree=tree, pk=pk)


def random_msa(N, L, rng, gap=0.1, other=0.0):
    """Unstructured random alignment (codes 0..4, plus N=15 with probability `other`)."""
    ax = rng.integers(0, 4, size=(N, L)).astype(np.uint8)
    ax[rng.random((N, L)) < gap] = 4
    if other > 0:
        ax[rng.random((N, L)) < other] = 15
    return ax
