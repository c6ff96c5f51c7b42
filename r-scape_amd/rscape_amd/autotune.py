"""Gram tile-shape autotuner.

Times every tile shape of the int8 one-hot Gram (rsg_gram_nshapes) on a given
alignment size and records the fastest per (N, L) bucket in a JSON table that
`best_shape` consults (tiles.json next to the library).  Usage:

    python -m rscape_amd.autotune --N 5000 --L 1500 [--reps 10] [--save]
"""
import argparse
import json
import os
import time

import numpy as np

from . import COUNTS_INT8, Context, lib, synth

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gram_tiles.json")


def bucket(N, L):
    """Size bucket: powers of two of N and L."""
    return "%d_%d" % (int(np.ceil(np.log2(max(N, 1)))), int(np.ceil(np.log2(max(L, 2)))))


def sweep(ctx, ax, reps=10):
    """Average Gram kernel time (ms) of every shape on alignment ax."""
    out = {}
    m = ctx.msa(ax)
    for s in range(1, lib().rsg_gram_nshapes() + 1):
        ctx.set_gram_shape(s)
        m.probs(COUNTS_INT8)                       # warm-up / JIT of the code object
        ctx.set_timing(True)
        for _ in range(reps):
            m.probs(COUNTS_INT8)
        t = [ms for name, ms in ctx.timings() if name == "gram_i8"]
        ctx.set_timing(False)
        out[s] = (lib().rsg_gram_shape_name(s).decode(), float(np.mean(t)))
    ctx.set_gram_shape(0)
    m.close()
    return out


def best_shape(N, L):
    """Recorded best shape for this size bucket, or 0 (the library default)."""
    try:
        return int(json.load(open(TABLE)).get(bucket(N, L), {}).get("shape", 0))
    except (OSError, ValueError):
        return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=5000)
    ap.add_argument("--L", type=int, default=1500)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--save", action="store_true")
    a = ap.parse_args()
    ax = synth.random_msa(a.N, a.L, np.random.default_rng(0), gap=0.1)
    ctx = Context(0)
    res = sweep(ctx, ax, a.reps)
    ops = 2 * 16 * a.N * a.L * (a.L - 1) / 2
    for s, (name, ms) in sorted(res.items(), key=lambda kv: kv[1][1]):
        print("shape %d %-20s %.4f ms  %.1f TOP/s" % (s, name, ms, ops / (ms * 1e-3) / 1e12))
    best = min(res, key=lambda s: res[s][1])
    if a.save:
        tab = {}
        if os.path.exists(TABLE):
            tab = json.load(open(TABLE))
        tab[bucket(a.N, a.L)] = {"shape": best, "name": res[best][0], "ms": res[best][1], "N": a.N, "L": a.L}
        json.dump(tab, open(TABLE, "w"), indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
