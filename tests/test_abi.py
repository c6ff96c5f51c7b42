"""The C-ABI library loads without a GPU and exports every entry point the header declares;
host-only entry points (no device needed) agree with the oracle."""
import ctypes as C

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
