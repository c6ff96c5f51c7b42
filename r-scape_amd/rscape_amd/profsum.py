"""Summaries of rocprofv3 CSV output (kernel statistics and per-dispatch counters).

    python -m rscape_amd.profsum stats <rocprof dir> [--top K]
    python -m rscape_amd.profsum traffic <FETCH_SIZE dir> <WRITE_SIZE dir> --kernel NAME
        --calib NAME --calib-read BYTES --calib-write BYTES
    python -m rscape_amd.profsum counters <pmc dir> --kernel NAME

`counters` prints the mean per dispatch of every counter collected for the kernels whose name
contains NAME.

`traffic` averages FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the kernel, scales each by
what the same counter reads for a calibration kernel of known bytes (known / measured), and
prints HBM bytes per launch.
"""
import argparse
import csv
import glob
import os
import sys


def _rows(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not files:
        sys.exit("no *%s under %s" % (suffix, d))
    for f in files:
        with open(f, newline="") as fh:
            yield from csv.DictReader(fh)


def stats(d, top):
    rows = list(_rows(d, "kernel_stats.csv"))
    rows.sort(key=lambda r: -float(r.get("TotalDurationNs", 0)))
    tot = sum(float(r.get("TotalDurationNs", 0)) for r in rows)
    out = ["%-60s %7s %12s %12s %7s" % ("kernel", "calls", "total_us", "avg_us", "pct")]
    for r in rows[:top]:
        out.append("%-60s %7s %12.1f %12.2f %6.1f%%" % (r["Name"][:60], r["Calls"], float(r["TotalDurationNs"]) / 1e3,
                                                        float(r["AverageNs"]) / 1e3,
                                                        100.0 * float(r["TotalDurationNs"]) / max(tot, 1.0)))
    return "\n".join(out)


def counter_avg(d, counter, name):
    vals = []
    for r in _rows(d, "counter_collection.csv"):
        if r.get("Counter_Name") == counter and name in r.get("Kernel_Name", ""):
            vals.append(float(r["Counter_Value"]))
    if not vals:
        sys.exit("no %s samples for kernel %s in %s" % (counter, name, d))
    return sum(vals) / len(vals), len(vals)


def counters(d, kernel):
    acc = {}
    for r in _rows(d, "counter_collection.csv"):
        if kernel in r.get("Kernel_Name", ""):
            acc.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    if not acc:
        sys.exit("no counters for kernel %s in %s" % (kernel, d))
    n = max(len(v) for v in acc.values())
    lines = ["kernel %s  (%d dispatches)" % (kernel, n)]
    for k in sorted(acc):
        lines.append("  %-28s %.4g" % (k, sum(acc[k]) / len(acc[k])))
    return "\n".join(lines)


def traffic(fd, wd, kernel, calib, cread, cwrite):
    kf, nf = counter_avg(fd, "FETCH_SIZE", kernel)
    kw, nw = counter_avg(wd, "WRITE_SIZE", kernel)
    cf, _ = counter_avg(fd, "FETCH_SIZE", calib)
    cw, _ = counter_avg(wd, "WRITE_SIZE", calib)
    sf, sw = cread / cf, cwrite / cw
    lines = [
        "kernel: %s  (%d FETCH dispatches, %d WRITE dispatches)" % (kernel, nf, nw),
        "FETCH_SIZE avg per dispatch: %.1f (counter units)" % kf,
        "WRITE_SIZE avg per dispatch: %.1f (counter units)" % kw,
        "calibration kernel: %s  reads %d B, writes %d B" % (calib, cread, cwrite),
        "calibration FETCH_SIZE %.1f -> scale %.3f B/unit; WRITE_SIZE %.1f -> scale %.3f B/unit" % (cf, sf, cw, sw),
        "fetch_bytes_per_launch: %.0f" % (kf * sf),
        "write_bytes_per_launch: %.0f" % (kw * sw),
        "traffic_bytes_per_launch: %.0f" % (kf * sf + kw * sw),
    ]
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats"); a.add_argument("dir"); a.add_argument("--top", type=int, default=25)
    b = sub.add_parser("traffic"); b.add_argument("fetch_dir"); b.add_argument("write_dir")
    b.add_argument("--kernel", required=True); b.add_argument("--calib", required=True)
    b.add_argument("--calib-read", type=float, required=True); b.add_argument("--calib-write", type=float, required=True)
    c = sub.add_parser("counters"); c.add_argument("dir"); c.add_argument("--kernel", required=True)
    args = ap.parse_args()
    if args.cmd == "stats":
        print(stats(args.dir, args.top))
    elif args.cmd == "counters":
        print(counters(args.dir, args.kernel))
    else:
        print(traffic(args.fetch_dir, args.write_dir, args.kernel, args.calib, args.calib_read, args.calib_write))


if __name__ == "__main__":
    main()
