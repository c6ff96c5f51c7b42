"""Gram tile-shape autotuner.

Times every tile shape of the int8 one-hot Gram (rsg_gram_nshapes) on alignments of given
sizes and records the fastest per size bucket (ceil(log2 N), ceil(log2 L)) in the table the
library reads on first use: gram_tiles.tsv next to librscape_hip.so, one "logN logL shape"
line per bucket.  Usage:

    python -m rscape_amd.autotune --N 5000 --L 1500 [--reps 10] [--save]
    python -m rscape_amd.autotune --grid [--save]      # every bucket of the default grid
    python -m rscape_amd.autotune --from-log sweep.log [more.log ...] --save
                                   # the table from the lines of earlier sweeps (later lines win)
"""
import argparse
import math
import os
import re

import numpy as np

from . import COUNTS_INT8, Context, lib, synth

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gram_tiles.tsv")
# bucket centres: N = 0.75 * 2^k for k = 9..15, L = 0.75 * 2^j for j = 8..13
GRID_N = [int(0.75 * 2 ** k) for k in range(9, 16)]
GRID_L = [int(0.75 * 2 ** j) for j in range(8, 14)]


def bucket(N, L):
    """Size bucket of the table: (ceil(log2 N), ceil(log2 L))."""
    return (int(math.ceil(math.log2(max(N, 1)))), int(math.ceil(math.log2(max(L, 2)))))


def sweep(ctx, ax, reps=10):
    """Average Gram time (ms) of every shape on alignment ax: {shape: (name, ms)}."""
    out = {}
    m = ctx.msa(ax)
    for s in range(1, lib().rsg_gram_nshapes() + 1):
        ctx.set_gram_shape(s)
        m.probs(COUNTS_INT8)                       # warm-up
        ctx.set_timing(True)
        for _ in range(reps):
            m.probs(COUNTS_INT8)
        t = [ms for name, ms in ctx.timings() if name == "gram_i8"]
        ctx.set_timing(False)
        out[s] = (lib().rsg_gram_shape_name(s).decode(), float(np.mean(t)))
    ctx.set_gram_shape(0)
    m.close()
    return out


def read_table(path=TABLE):
    tab = {}
    if os.path.exists(path):
        for line in open(path):
            line = line.split("#", 1)[0].split()
            if len(line) == 3:
                tab[(int(line[0]), int(line[1]))] = int(line[2])
    return tab


def write_table(tab, path=TABLE, notes=None):
    with open(path, "w") as f:
        f.write("# int8 Gram tile shape per size bucket: ceil(log2 N) ceil(log2 L) shape (rsg_gram_shape_name)\n")
        f.write("# written by python -m rscape_amd.autotune on an MI355X; sizes without a line use the built-in rule\n")
        for (a, b), s in sorted(tab.items()):
            f.write("%d %d %d%s\n" % (a, b, s, ("  # " + notes[(a, b)]) if notes and (a, b) in notes else ""))


_LINE = re.compile(r"N=(\d+) L=(\d+) best (\d+) (\S+) ([\d.]+) ms \((\d+) TOP/s\); built-in choice ([\d.]+) ms")


def from_logs(paths):
    """Table entries and notes from the result lines a sweep printed (later lines win)."""
    tab, notes = {}, {}
    for path in paths:
        for line in open(path):
            m = _LINE.search(line)
            if not m:
                continue
            N, L, best = int(m.group(1)), int(m.group(2)), int(m.group(3))
            tab[bucket(N, L)] = best
            notes[bucket(N, L)] = "%s, N=%d L=%d: %s ms (built-in choice %s ms)" % (m.group(4), N, L, m.group(5), m.group(7))
    return tab, notes


def best_shape(N, L):
    """Recorded best shape for this size bucket, or 0 (the library default)."""
    return read_table().get(bucket(N, L), 0)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=5000)
    ap.add_argument("--L", type=int, default=1500)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--grid", action="store_true", help="sweep every bucket of the default grid")
    ap.add_argument("--save", action="store_true")
    ap.add_argument("--out", default=TABLE, help="table to update with --save")
    ap.add_argument("--from-log", nargs="+", default=None, help="build the table from printed sweep lines")
    a = ap.parse_args()
    if a.from_log:
        tab, notes = from_logs(a.from_log)
        print("%d buckets from %s" % (len(tab), ", ".join(a.from_log)))
        if a.save:
            write_table(tab, a.out, notes=notes)
        return
    sizes = [(n, l) for n in GRID_N for l in GRID_L] if a.grid else [(a.N, a.L)]
    ctx = Context(0)
    tab, notes = read_table(a.out), {}
    for N, L in sizes:
        ax = synth.random_msa(N, L, np.random.default_rng(N + L), gap=0.1)
        res = sweep(ctx, ax, a.reps)
        ops = 2 * 16 * N * L * (L - 1) / 2
        best = min(res, key=lambda s: res[s][1])
        dflt = res[7 if N >= 1024 else 4][1]
        print("N=%d L=%d best %d %s %.4f ms (%.0f TOP/s); built-in choice %.4f ms" %
              (N, L, best, res[best][0], res[best][1], ops / (res[best][1] * 1e-3) / 1e12, dflt), flush=True)
        tab[bucket(N, L)] = best
        notes[bucket(N, L)] = "%s, N=%d L=%d: %.4f ms (built-in choice %.4f ms)" % (res[best][0], N, L, res[best][1], dflt)
    if a.save:
        write_table(tab, a.out, notes=notes)
    ctx.close()


if __name__ == "__main__":
    main()
