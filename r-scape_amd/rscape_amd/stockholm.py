"""Minimal Stockholm reader with Easel's RNA digitisation.

Digital codes follow Easel's RNA alphabet "ACGU-RYMKSWHBVDN*~" (0..17); T reads as U,
X as N, and '.', '_', '=' as gap (R-scape adds '=' and '.' as gap characters,
src/R-scape.c:1103-1104).  Lower case reads as upper case.
"""
import numpy as np

SYMBOLS = "ACGU-RYMKSWHBVDN*~"
_MAP = np.full(256, 255, dtype=np.uint8)
for _k, _c in enumerate(SYMBOLS):
    _MAP[ord(_c)] = _k
    _MAP[ord(_c.lower())] = _k
for _c, _t in (("T", "U"), ("X", "N"), ("I", "A"), (".", "-"), ("_", "-"), ("=", "-")):
    _MAP[ord(_c)] = _MAP[ord(_t)]
    _MAP[ord(_c.lower())] = _MAP[ord(_t)]


def digitize(s):
    a = _MAP[np.frombuffer(s.encode("ascii"), dtype=np.uint8)]
    if (a == 255).any():
        bad = s[int(np.argmax(a == 255))]
        raise ValueError("illegal character %r in alignment" % bad)
    return a


def read(path):
    """Read the first alignment of a Stockholm file: (names, ax (N x L uint8), gc dict)."""
    seqs, order, gc = {}, [], {}
    with open(path) as fh:
        for line in fh:
            if line.startswith("//"):
                break
            if line.startswith("#=GC"):
                parts = line.split()
                if len(parts) >= 3:
                    gc[parts[1]] = gc.get(parts[1], "") + parts[2]
                continue
            if line.startswith("#") or not line.strip():
                continue
            parts = line.split()
            if len(parts) < 2:
                continue
            n, s = parts[0], parts[1]
            if n not in seqs:
                seqs[n] = []
                order.append(n)
            seqs[n].append(s)
    rows = ["".join(seqs[n]) for n in order]
    if len({len(r) for r in rows}) != 1:
        raise ValueError("sequences of unequal aligned length")
    ax = np.stack([digitize(r) for r in rows])
    return order, ax, gc


def write(path, names, ax):
    with open(path, "w") as fh:
        fh.write("# STOCKHOLM 1.0\n\n")
        w = max(len(n) for n in names)
        for n, row in zip(names, ax):
            fh.write("%-*s %s\n" % (w, n, "".join(SYMBOLS[c] for c in row)))
        fh.write("//\n")
