"""Summaries of rocprofv3 output for the grid kernels (markdown text).

  python -m lcpao_grid.profsum stats  KERNEL_STATS_CSV [--title T] [--command C]
  python -m lcpao_grid.profsum traffic --fetch FETCH_CSV --write WRITE_CSV
         --copy-match SUBSTR --copy-bytes N [--title T] [--command C]

`stats` condenses a `--kernel-trace --stats` run (kernel_stats.csv) to a
table of calls and average/total durations.  `traffic` turns two `--pmc`
runs (FETCH_SIZE, then WRITE_SIZE; both in KiB per dispatch) into HBM bytes
per dispatch of every kernel, each counter scaled by what it reads for a
plain copy kernel that moves a known number of bytes (N read and N written):
factor = N / (counter KiB * 1024).  The table that `traffic` prints is read
back by bench.py (column "traffic B/dispatch").
"""
import argparse
import csv
import re
import sys
from collections import defaultdict


def short_name(name):
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    depth, out = 0, []
    for ch in name:  # drop the argument list, keep template arguments
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            break
        out.append(ch)
    s = "".join(out).strip()
    return re.sub(r"\s+", " ", s)


def read_stats(path):
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((short_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]),
                         float(r["AverageNs"]), float(r.get("Percentage", 0.0) or 0.0)))
    rows.sort(key=lambda x: -x[2])
    return rows


def read_counter(path, counter):
    per = defaultdict(list)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            per[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (len(v), sum(v) / len(v)) for k, v in per.items()}


def cmd_stats(a):
    rows = read_stats(a.csv)
    print("# %s" % (a.title or "rocprofv3 kernel statistics"))
    print()
    if a.command:
        print("Command: `%s`" % a.command)
        print()
    print("| kernel | calls | avg ms | total ms | % |")
    print("|---|---:|---:|---:|---:|")
    for name, calls, tot, avg, pct in rows:
        print("| `%s` | %d | %.4f | %.3f | %.2f |" % (name, calls, avg * 1e-6, tot * 1e-6, pct))


def cmd_traffic(a):
    fetch = read_counter(a.fetch, "FETCH_SIZE")
    write = read_counter(a.write, "WRITE_SIZE")
    cf = [k for k in fetch if a.copy_match in k]
    cw = [k for k in write if a.copy_match in k]
    if not cf or not cw:
        sys.exit("copy kernel matching %r not found" % a.copy_match)
    f_copy, w_copy = fetch[cf[0]][1], write[cw[0]][1]
    ff = a.copy_bytes / (f_copy * 1024.0)
    fw = a.copy_bytes / (w_copy * 1024.0)
    print("# %s" % (a.title or "HBM traffic per dispatch (rocprofv3 --pmc)"))
    print()
    if a.command:
        print("Commands: `%s`" % a.command)
        print()
    print("Calibration: copy kernel `%s` moves %d bytes in and %d bytes out per dispatch; "
          "FETCH_SIZE reads %.1f KiB (factor %.4f), WRITE_SIZE reads %.1f KiB (factor %.4f)."
          % (cf[0], a.copy_bytes, a.copy_bytes, f_copy, ff, w_copy, fw))
    print()
    print("| kernel | dispatches | FETCH_SIZE KiB | WRITE_SIZE KiB | fetch B | write B | traffic B/dispatch |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    names = sorted(set(fetch) & set(write), key=lambda k: -(fetch[k][1] + write[k][1]))
    for k in names:
        fb = fetch[k][1] * 1024.0 * ff
        wb = write[k][1] * 1024.0 * fw
        print("| `%s` | %d | %.1f | %.1f | %.4g | %.4g | %.6g |" % (k, fetch[k][0], fetch[k][1], write[k][1], fb, wb, fb + wb))


def read_traffic_table(path):
    """{kernel: traffic bytes per dispatch} from a `traffic` summary."""
    out = {}
    with open(path) as f:
        for line in f:
            m = re.match(r"\|\s*`([^`]+)`\s*\|(.*)\|\s*$", line)
            if not m:
                continue
            cols = [c.strip() for c in m.group(2).split("|")]
            if len(cols) == 6:
                out[m.group(1)] = float(cols[5])
    return out


def main(argv=None):
    p = argparse.ArgumentParser(prog="lcpao_grid.profsum")
    sub = p.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--title", default="")
    s.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-match", required=True)
    t.add_argument("--copy-bytes", type=float, required=True)
    t.add_argument("--title", default="")
    t.add_argument("--command", default="")
    a = p.parse_args(argv)
    {"stats": cmd_stats, "traffic": cmd_traffic}[a.cmd](a)


if __name__ == "__main__":
    main()
