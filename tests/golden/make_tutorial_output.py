"""Writes tests/golden/arisong_output.json: the significant-pair lines (two-set test) and the
power block that documentation/tutorial.tex:193-235 publishes for tutorial/updated_Arisong.sto,
split into whitespace-separated tokens (the document shows the files' tabs as spaces).

    python tests/golden/make_tutorial_output.py /root/reference/documentation/tutorial.tex
"""
import json
import os
import sys


def main(tex):
    lines = open(tex).read().splitlines()
    start = next(k for k, l in enumerate(lines) if "bin/R-scape -s tutorial" in l)
    block = []
    for l in lines[start:]:
        if l.startswith("# RMs"):
            break
        block.append(l)
    head = next(l for l in block if l.startswith("# GTp"))
    hits = [l.split() for l in block if l.startswith("*")]
    p0 = next(k for k, l in enumerate(block) if l.startswith("# Powertype"))
    power = [l for l in block[p0:] if l.strip()]
    out = {"source": "documentation/tutorial.tex:193-235 (R-scape -s tutorial/updated_Arisong.sto)",
           "summary": head.split(), "hits": hits, "power_block": [l.split() for l in power]}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "arisong_output.json")
    json.dump(out, open(path, "w"), indent=1)
    print("wrote", path, len(hits), "hits", len(power), "power lines")


if __name__ == "__main__":
    main(sys.argv[1])
