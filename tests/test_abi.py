"""The C-ABI library loads without a GPU and exports every entry point the header declares;
host-only entry points (no device needed) agree with the oracle."""
import ctypes as C
import os

import numpy as np

import oracle
import rscape_amd as R


def test_library_exports_every_declared_symbol():
    names = R.exported_symbols()
    assert len(names) > 40
    L = R.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_gapcol_mask_matches_oracle():
    rng = np.random.default_rng(0)
    for N, L, thr in [(1, 1, 0.75), (30, 50, 0.75), (200, 120, 0.5), (57, 300, 0.0), (57, 300, 1.0)]:
        ax = rng.integers(0, 4, size=(N, L), dtype=np.uint8)
        ax[rng.random((N, L)) < rng.random(L)[None, :]] = 4
        ax[rng.random((N, L)) < 0.02] = 17
        ax[rng.random((N, L)) < 0.02] = 9
        ax[rng.random((N, L)) < 0.01] = 16
        w = rng.random(N) + 0.1
        use, kept = R.gapcol_mask(ax, w, thr)
        ref = oracle.gapcol_filter(ax, w, thr)
        assert np.array_equal(use, ref) and kept == ref.sum()


def test_host_histogram_entry_points_match_oracle():
    """rsg_width and rsg_cov2evalue are host-side FP64; compare with the oracle."""
    for mx, mn in [(50.0, -3.0), (0.5, 0.1), (1e-7, 0.0), (300.0, -50.0)]:
        assert R.width(mx, mn) == oracle.lib().ref_width(mx, mn, -10.0, 0.05, 400, 1e-6)


def _lib_hist(bmin, bmax, w, counts):
    """A library histogram (rsg_hist_create) with the given bin counts."""
    L = R.lib()
    p = L.rsg_hist_create(bmin, bmax, w)
    h = p.contents
    assert h.nb == len(counts)
    for b, c in enumerate(counts):
        h.obs[b] = int(c)
    nz = np.nonzero(counts)[0]
    h.imin, h.imax = int(nz[0]), int(nz[-1])
    h.n = h.Nc = h.No = int(np.sum(counts))
    return R.Histogram(p)


def _fields(h):
    return (h.nb, h.bmin, h.bmax, h.w, h.imin, h.imax, h.cmin, h.n, h.Nc, h.No, h.xmin, h.xmax)


def test_null_file_format_and_read_back(tmp_path):
    """cov_WriteNullHistogram / cov_ReadNullHistogram (covariation.c:1717-1868): one "%f %llu"
    line per bin (its lower bound); reading back infers w as the smallest score gap and puts each
    line's count into Score2Bin(score).  A lower bound lies on the boundary of the bin below, so
    with exactly representable scores (w = 1/8) every count moves down one bin and the histogram
    grows two bins below; the library's reader equals the oracle's restatement, also for w = 0.05
    (whose printed scores round) and for lines that set no bin (one or three tokens)."""
    rng = np.random.default_rng(4)
    nb, w, bmin = 40, 0.125, -10.0
    counts = rng.integers(0, 1000, size=nb).astype(np.uint64)
    counts[[0, nb - 1]] = [3, 7]
    h = _lib_hist(bmin, bmin + nb * w, w, counts)
    f = str(tmp_path / "a.null")
    h.write_null(f)
    lines = open(f).read().splitlines()
    assert lines == ["%f %d" % (bmin + i * w, c) for i, c in enumerate(counts)]
    gh = R.Histogram.read_null(f)                 # keeps the library histogram alive
    g = gh.h
    assert g.w == w and g.nb == nb + 1 and g.bmin == bmin - 2 * w
    got = g.counts()
    assert got[0] == 0 and np.array_equal(got[1:], counts)       # bin i + 2 - 1: one bin lower
    assert (g.imin, g.imax, g.n) == (1, nb, int(counts.sum()))
    r = oracle._hp(oracle.read_null(f)).contents
    assert _fields(g) == _fields(r) and np.array_equal(got, r.counts())
    # rounded scores and malformed lines: library reader = oracle reader, field by field
    counts = rng.integers(0, 50, size=300).astype(np.uint64)
    h = _lib_hist(-10.0, -10.0 + 300 * 0.05, 0.05, counts)
    f2 = str(tmp_path / "b.null")
    h.write_null(f2)
    with open(f2, "a") as fh:
        fh.write("12.5\n3.0 4 5\n")
    gh2 = R.Histogram.read_null(f2)
    g, r = gh2.h, oracle._hp(oracle.read_null(f2)).contents
    assert _fields(g) == _fields(r) and np.array_equal(g.counts(), r.counts())
    assert int(g.n) == int(counts.sum())


def test_gram_tiles_table_load(tmp_path):
    """The autotuner's table loads without a device; malformed lines are rejected with the line
    number; the committed table (read by the library on first use) parses."""
    from rscape_amd import autotune
    L = R.lib()
    err = C.create_string_buffer(256)
    bad = tmp_path / "bad.tsv"
    bad.write_text("# comment\n13 11 7\n13 11\n")
    assert L.rsg_gram_tiles_load(str(bad).encode(), err) != 0 and b"line 3" in err.value
    bad.write_text("13 11 99\n")
    assert L.rsg_gram_tiles_load(str(bad).encode(), err) != 0
    assert L.rsg_gram_tiles_load(str(tmp_path / "missing.tsv").encode(), err) != 0
    good = tmp_path / "good.tsv"
    autotune.write_table({(13, 11): 7, (10, 9): 4}, str(good))
    assert L.rsg_gram_tiles_load(str(good).encode(), err) == 0
    assert autotune.read_table(str(good)) == {(13, 11): 7, (10, 9): 4}
    if os.path.exists(autotune.TABLE):                  # back to the committed table
        assert L.rsg_gram_tiles_load(autotune.TABLE.encode(), err) == 0, err.value
        tab = autotune.read_table()
        assert tab and all(1 <= v <= L.rsg_gram_nshapes() for v in tab.values())
        # the committed table covers every bucket of the default sweep grid
        assert {autotune.bucket(n, l) for n in autotune.GRID_N for l in autotune.GRID_L} <= set(tab)


def test_autotune_table_from_sweep_lines(tmp_path):
    """The table rebuilt from printed sweep lines: one entry per bucket, later lines win, other
    lines ignored."""
    from rscape_amd import autotune
    a, b = tmp_path / "a.log", tmp_path / "b.log"
    a.write_text("[sweep] unrelated line\n"
                 "N=6144 L=1536 best 8 32x64c/4w/kt256 0.0933 ms (2484 TOP/s); built-in choice 0.0949 ms\n"
                 "N=384 L=192 best 8 32x64c/4w/kt256 0.0123 ms (44 TOP/s); built-in choice 0.0130 ms\n")
    b.write_text("N=5000 L=1500 best 7 64x32c/4w/kt256 0.0744 ms (2418 TOP/s); built-in choice 0.0744 ms\n")
    tab, notes = autotune.from_logs([str(a), str(b)])
    assert tab == {(13, 11): 7, (9, 8): 8}
    assert "N=5000 L=1500" in notes[(13, 11)]
    out = tmp_path / "t.tsv"
    autotune.write_table(tab, str(out), notes=notes)
    assert autotune.read_table(str(out)) == tab
