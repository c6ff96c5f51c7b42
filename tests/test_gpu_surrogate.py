"""Surrogate inputs and outputs on the device: the 64 features against a numpy statement of
their definition, the batch's per-group features against the single-alignment entry point,
the network's scalers, and tail E-values against numpy."""
import numpy as np
import pytest

import rscape_amd as R
from rscape_amd import null as NL
from rscape_amd import synth
from rscape_amd.mlp import MLP

pytestmark = pytest.mark.gpu


def features_np(ax, tree):
    ax = np.asarray(ax)
    N, L = ax.shape
    left, right, ld, rd = (np.asarray(tree[k]) for k in ("left", "right", "ld", "rd"))
    f = np.zeros(64)
    bl = np.concatenate([ld, rd])
    T = bl.sum()
    f[0], f[1], f[2] = np.log(N), np.log(L), np.log(T)
    cnt = np.stack([(ax == x).sum(0) for x in range(4)], 0).astype(np.float64)   # 4 x L
    nc = cnt.sum(0)
    with np.errstate(divide="ignore", invalid="ignore"):
        p = np.where(nc > 0, cnt / nc, 0.0)
        H = -np.sum(np.where(p > 0, p * np.log2(np.where(p > 0, p, 1.0)), 0.0), 0)
    g = (N - nc) / N
    f[3:19] = np.bincount(np.minimum(15, (H * 8).astype(int)), minlength=16) / L
    f[19:35] = np.bincount(np.minimum(15, (g * 16).astype(int)), minlength=16) / L
    f[35:38] = cnt.sum(1)[:3] / cnt.sum()
    mb = T / len(bl)
    f[38], f[39] = mb, bl.std() / mb
    depth = np.zeros(N - 1); edges = np.zeros(N - 1)
    leafd, leafe = np.zeros(N), np.zeros(N)
    stack = [0]
    while stack:
        v = stack.pop()
        for d, b in ((left[v], ld[v]), (right[v], rd[v])):
            if d > 0:
                depth[d], edges[d] = depth[v] + b, edges[v] + 1
                stack.append(d)
            else:
                leafd[-d], leafe[-d] = depth[v] + b, edges[v] + 1
    md = leafd.mean()
    f[40], f[41], f[42] = md, leafd.std() / md, leafd.max() / md
    term = np.concatenate([ld[left <= 0], rd[right <= 0]])
    inter = np.concatenate([ld[left > 0], rd[right > 0]])
    f[43] = term.mean() / mb
    f[44] = inter.mean() / mb if len(inter) else 0.0
    f[45] = leafe.mean() / np.log2(N)
    f[46] = np.sum((left <= 0) & (right <= 0)) / (N / 2)
    nleaf = np.zeros(N - 1)
    coll = 0.0
    for v in range(N - 2, -1, -1):          # children have larger indices in these trees
        a = nleaf[left[v]] if left[v] > 0 else 1
        b = nleaf[right[v]] if right[v] > 0 else 1
        nleaf[v] = a + b
        coll += abs(a - b)
    f[47] = coll / ((N - 1) * (N - 2) / 2) if N > 2 else 0.0
    # split spectrum: branch sums by the minor clade share q of the branch's split,
    # bins q in (1/4, 1/2], (1/8, 1/4], ..., (1/256, 1/128], <= 1/256
    s1, s2 = np.zeros(8), np.zeros(8)
    for v in range(N - 1):
        for d, t in ((left[v], ld[v]), (right[v], rd[v])):
            n = nleaf[d] if d > 0 else 1
            q = min(n, N - n) / N
            k = min(7, int(np.floor(-np.log2(q))) - 1)          # q in (2^-(k+2), 2^-(k+1)]
            s1[k] += t
            s2[k] += (1 - np.exp(-t)) ** 2
    f[48:56] = np.log(1e-3 + s1)
    f[56:64] = np.log(1e-4 + s2)
    return f


@pytest.mark.parametrize("N,L,gap", [(3, 5, 0.0), (200, 120, 0.05), (1000, 700, 0.3)])
def test_features_match_numpy(ctx, N, L, gap):
    m = NL.gtr_model((0.3, 0.2, 0.2, 0.3))
    tree = NL.yule_tree(N, 0.1, seed=N)
    leaves = NL.simulate(ctx, NL.Job(tree, L, m, seed=3, gapfrac=gap))
    msa = ctx.msa(leaves)
    got = NL.features(ctx, msa, tree)
    ref = features_np(leaves, tree)
    assert np.allclose(got, ref, rtol=2e-6, atol=1e-6), np.nonzero(~np.isclose(got, ref, rtol=2e-6, atol=1e-6))


def test_batch_features_equal_single(ctx):
    m = NL.gtr_model((0.25, 0.22, 0.30, 0.23))
    trees = [NL.yule_tree(n, 0.1 + 0.05 * k, seed=k) for k, n in enumerate([30, 300, 77])]
    jobs = []
    for g, (t, L) in enumerate(zip(trees, [64, 200, 129])):
        jobs += [NL.Job(t, L, m, seed=11, stream=r, gapfrac=0.1, group=g) for r in range(2)]
    fits, _, fe = NL.batch(ctx, jobs, "GTp", pmass=0.05, feats=True)
    for g in range(3):
        job = jobs[2 * g]
        single = NL.features(ctx, ctx.msa(NL.simulate(ctx, job)), job.tree)
        assert np.array_equal(fe[g], single)


def test_mlp_scaling(ctx):
    rng = np.random.default_rng(2)
    dims = (48, 64, 2)
    a = MLP(ctx, dims, seed=5, bf16=False)
    b = MLP(ctx, dims, seed=5, bf16=False)
    xm, xs = rng.normal(size=48), rng.uniform(0.5, 2, 48)
    ym, ys = np.array([3.0, -1.0]), np.array([2.0, 0.5])
    b.set_scaling(xm, xs, ym, ys)
    x = rng.normal(size=(70, 48)).astype(np.float32)
    ya = a.forward(((x - xm) / xs).astype(np.float32))
    yb = b.forward(x)
    assert np.allclose(yb, ya * ys + ym, rtol=1e-5, atol=1e-5)
    # training targets are scaled the same way: gradients of b on raw data = a on scaled data
    t = rng.normal(size=(70, 2)).astype(np.float32)
    ga = a.grad(((x - xm) / xs).astype(np.float32), ((t - ym) / ys).astype(np.float32))
    gb = b.grad(x, t)
    for u, v in zip(ga[0], gb[0]):
        assert np.allclose(u, v, rtol=1e-4, atol=1e-6)
    b.set_range(np.full(48, -1.0), np.full(48, 1.0))
    assert b.in_range(np.zeros(48)) and not b.in_range(np.full(48, 2.0))
    b.set_range()
    assert b.in_range(np.full(48, 2.0))


def test_evalues_tail(ctx):
    d = synth.config(0, seed=4, N=120, L=60)
    msa = ctx.msa(d["ax"])
    msa.calculate("GTp")
    cov, _, _ = msa.get_cov()
    ev = NL.evalues_tail(ctx, msa, 5.0, 0.4, 1e-3)
    Nc = 60 * 59 // 2
    ref = Nc * np.minimum(1.0, 1e-3 * np.exp(-0.4 * (cov - 5.0)))
    off = ~np.eye(60, dtype=bool)
    assert np.allclose(ev[off], ref[off], rtol=1e-12)
    assert np.all(np.isinf(np.diag(ev)))
    # a cap inside the score range: every score at or above it gets the E-value at the cap
    xcap = float(np.percentile(cov[off], 99))
    evc = NL.evalues_tail(ctx, msa, 5.0, 0.4, 1e-3, xcap=xcap)
    refc = Nc * np.minimum(1.0, 1e-3 * np.exp(-0.4 * (np.minimum(cov, xcap) - 5.0)))
    assert np.allclose(evc[off], refc[off], rtol=1e-12)
    top = off & (cov >= xcap)
    assert top.sum() >= 10 and np.all(evc[top] == evc[top][0])



@pytest.mark.parametrize("thresh", [1e-3, 1e300])
def test_hit_list_matches_evalue_matrix(ctx, thresh):
    """rsg_cov_Hits: the pairs i < j with E <= thresh, sorted by (i, j), with their scores and
    E-values, as read off the full matrices on the host -- a few hits, and every pair (more than
    the 4096 fetched with the count, so the remainder takes a second copy); called twice in a
    row (the compaction buffers are reused)."""
    L = 120
    d = synth.config(0, seed=6, N=150, L=L)
    msa = ctx.msa(d["ax"])
    msa.calculate("GTp")
    cov, _, _ = msa.get_cov()
    ev = NL.evalues_tail(ctx, msa, 5.0, 0.4, 1e-3)
    iu = np.triu_indices(L, 1)
    sel = ev[iu] <= thresh
    ri, rj = iu[0][sel], iu[1][sel]
    assert 0 < len(ri) and (thresh < 1 or len(ri) == L * (L - 1) // 2 > 4096)
    for _ in range(2):
        NL.evalues_tail(ctx, msa, 5.0, 0.4, 1e-3, matrix=False)
        hi, hj, sc, he = NL.hits(ctx, msa, thresh)
        assert np.array_equal(hi, ri) and np.array_equal(hj, rj)
        assert np.array_equal(sc, cov[ri, rj]) and np.array_equal(he, ev[ri, rj])
    msa.close()

def test_surrogate_beats_constant_tail(ctx):
    """Trained on batched-null fits of small random setups, the network predicts held-out
    tails better than the method without it (one tail for every alignment)."""
    from rscape_amd import surrogate as S
    X, Y, _ = S.generate(ctx, 400, seed=3, Nr=(20, 300), Lr=(50, 160), budget=4e9)
    assert len(X) > 300 and np.all(np.isfinite(X)) and np.all(np.isfinite(Y))
    rng = np.random.default_rng(0)
    idx = rng.permutation(len(X))
    va, tr = idx[:80], idx[80:]
    sur = S.Surrogate(ctx, seed=2)
    sur.fit(X[tr], Y[tr], epochs=200, batch=64, lr=2e-3, seed=5)
    pm, pl = sur.predict(X[va])
    ok = np.isfinite(pm)
    assert ok.mean() > 0.7
    # log10 tail probability error 5 and 15 e-folds into each held-out tail
    lt, ut = np.exp(Y[va, 1][ok]), Y[va, 0][ok]             # true lambda, lambda * m
    ls, us = pl[ok], pm[ok] * pl[ok]
    lc, uc = np.exp(Y[tr, 1].mean()), Y[tr, 0].mean()        # one tail for every alignment
    errs, cerrs = [], []
    for k in (5.0, 15.0):
        x = (ut + k) / lt                                    # lambda_t (x - m_t) = k
        errs.append(np.abs(ls * x - us - k) / np.log(10))
        cerrs.append(np.abs(lc * x - uc - k) / np.log(10))
    e, ce = np.median(np.concatenate(errs)), np.median(np.concatenate(cerrs))
    print("median log10 tail error: network %.3f, constant tail %.3f" % (e, ce))
    assert e < 0.5 * ce


def test_full_nulls_batch_equals_single_nulls(ctx):
    """Several independent nulls of one alignment in one batch (one group each) = each null on
    its own (same seeds and streams): same histograms and fits."""
    from rscape_amd import null as NL
    from rscape_amd import surrogate as S
    tree = NL.yule_tree(60, 0.1, seed=3)
    pi, gap = (0.3, 0.2, 0.2, 0.3), 0.05
    many = S.full_nulls(ctx, tree, 90, pi, gap, [11, 12, 13], K=4)
    for s, (h, sf, fit) in zip([11, 12, 13], many):
        h1, sf1, fit1 = S.full_null(ctx, tree, 90, pi, gap, K=4, seed=s)
        assert np.array_equal(h.h.counts(), h1.h.counts())
        assert np.array_equal(np.nan_to_num(fit), np.nan_to_num(fit1))
        assert np.array_equal(np.nan_to_num(sf), np.nan_to_num(sf1))


def test_surrogate_cap_network_and_save_load(ctx, tmp_path):
    """The cap network learns each group's offset of the largest null score from the linear cap
    model from the inputs and the null size, a saved and reloaded surrogate gives the same tail
    and cap, and tail() takes the offset at the null's own size (clamped to the training range)."""
    from rscape_amd import surrogate as S
    X, Y, st = S.generate(ctx, 300, seed=4, Nr=(20, 300), Lr=(50, 160), budget=4e9, target_pairs=(2e4, 4e5))
    assert X.shape[1] == S.DIMS[0] and np.all(np.isfinite(X))
    sur = S.Surrogate(ctx, seed=2)
    sur.fit(X, Y, epochs=20, batch=64, lr=2e-3, seed=5)
    a, b = sur.fit_cap(st, Y, X=X, epochs=10, batch=64, lr=1e-3)
    z, u = S.Surrogate.zu(st, Y)
    ok = np.isfinite(z)
    assert abs(np.mean(z[ok] - (a + b * u[ok]))) < 1e-6           # least-squares line
    assert len(set(s["K"] for s in st)) > 3                           # groups of several sizes
    dl = sur.delta(X[:16], 5.0)
    assert dl.shape == (16,) and np.all(np.isfinite(dl))
    lo, hi = sur.cap_urange
    assert np.array_equal(sur.delta(X[:4], hi + 5.0), sur.delta(X[:4], hi))     # u is clamped
    path = str(tmp_path / "sur.npz")
    sur.save(path)
    back = S.Surrogate.load(ctx, path)
    for k in range(4):
        t1, t2 = sur.tail(X[k], 20, 5000), back.tail(X[k], 20, 5000)
        assert np.allclose(t1, t2, rtol=1e-6, equal_nan=True)
        if np.isfinite(t1[0]):
            m, lam, _ = t1
            dl = float(sur.delta(X[k], np.log(20 * 5000 * S.P0))[0])
            assert t1[2] == sur.cap(m, lam, 20, 5000, delta=dl)
