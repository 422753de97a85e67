"""Summaries of rocprofv3 output for the grid kernels (markdown text).

  python profsum.py stats  KERNEL_STATS_CSV [--title T] [--command C]
  python profsum.py traffic --fetch FETCH_CSV --write WRITE_CSV
         --copy-match SUBSTR --copy-bytes N [--title T] [--command C]
  python profsum.py mfma COUNTER_CSV --agent AGENT_INFO_CSV
         [--title T] [--command C]
  python profsum.py counters COUNTER_CSV [--match SUBSTR ...]
         [--title T] [--command C]

`stats` condenses a `--kernel-trace --stats` run (kernel_stats.csv) to a
table of calls and average/total durations.  `traffic` turns two `--pmc`
runs (FETCH_SIZE, then WRITE_SIZE; both in KiB per dispatch) into HBM bytes
per dispatch of every kernel, each counter scaled by what it reads for a
plain copy kernel that moves a known number of bytes (N read and N written):
factor = N / (counter KiB * 1024).  The table that `traffic` prints is read
back by bench.py (column "traffic B/dispatch").

`counters` prints the per-dispatch average of every counter of one `--pmc`
run for the kernels whose names contain one of the --match strings, with
each SQ_* wave-cycle counter also as a fraction of SQ_WAVE_CYCLES when that
counter is present.

`mfma` turns one `--pmc` run of SQ_INSTS_VALU_MFMA_MOPS_{BF16,F32,F64},
SQ_VALU_MFMA_BUSY_CYCLES and GRBM_GUI_ACTIVE into the matrix-core use of
every kernel: executed MFMA FLOPs per dispatch (512 per MOP, rocprofv3's own
derived-counter rule), their rate over the dispatch's duration against the
MI355X dense peak of that type, and the busy fraction
SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE per XCC * SIMDs) (rocprofv3's
MFMA-utilisation expression; the collected GRBM_GUI_ACTIVE is the sum over
the XCCs, so it is divided by their number from the agent info).  Durations
come from the counter run's own timestamps (serialised dispatches, slightly
longer than in a plain run); bench.py reads the busy fraction back.
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
    """{kernel: (dispatches, traffic bytes per dispatch)} from a `traffic` summary."""
    out = {}
    with open(path) as f:
        for line in f:
            m = re.match(r"\|\s*`([^`]+)`\s*\|(.*)\|\s*$", line)
            if not m:
                continue
            cols = [c.strip() for c in m.group(2).split("|")]
            if len(cols) == 6:
                out[m.group(1)] = (int(cols[0]), float(cols[5]))
    return out


MFMA_PEAK_TF = {"BF16": 2500.0, "F32": 157.3, "F64": 78.6}  # MI355X dense matrix peaks


def read_mfma(path):
    """{kernel: dict(n, ms, mops{type}, busy, gui)} averaged over dispatches."""
    acc = defaultdict(lambda: defaultdict(list))
    times = defaultdict(dict)
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            k = short_name(r["Kernel_Name"])
            acc[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
            times[k][r["Dispatch_Id"]] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    out = {}
    for k, c in acc.items():
        avg = {n: sum(v) / len(v) for n, v in c.items()}
        ts = list(times[k].values())
        out[k] = {"n": len(ts), "ms": sum(ts) / len(ts) * 1e-6, "busy": avg.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0),
                  "gui": avg.get("GRBM_GUI_ACTIVE", 0.0),
                  "mops": {t: avg.get("SQ_INSTS_VALU_MFMA_MOPS_" + t, 0.0) for t in MFMA_PEAK_TF}}
    return out


def read_agent(path):
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Agent_Type"] == "GPU":
                return int(r["Simd_Count"]), int(r["Num_Xcc"]), r["Gfx_Target_Version"]
    sys.exit("no GPU agent in %s" % path)


def cmd_mfma(a):
    simds, nxcc, gfx = read_agent(a.agent)
    rows = read_mfma(a.csv)
    print("# %s" % (a.title or "Matrix-core use per dispatch (rocprofv3 --pmc)"))
    print()
    if a.command:
        print("Command: `%s`" % a.command)
        print()
    print("Agent: %d SIMDs in %d XCCs (gfx target %s).  Executed MFMA FLOPs = 512 x SQ_INSTS_VALU_MFMA_MOPS_<type>; "
          "rate over the counter run's dispatch duration; peak = MI355X dense matrix peak of the type (BF16 2500, "
          "F32 157.3, F64 78.6 TFLOP/s); busy = SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / %d x %d)."
          % (simds, nxcc, gfx, nxcc, simds))
    print()
    print("| kernel | dispatches | ms | type | executed GFLOP/dispatch | MFMA TFLOP/s | of type peak | MFMA busy |")
    print("|---|---:|---:|---|---:|---:|---:|---:|")
    for k in sorted(rows, key=lambda x: -max(rows[x]["mops"].values())):
        r = rows[k]
        typ = max(r["mops"], key=lambda t: r["mops"][t])
        if r["mops"][typ] <= 0:
            continue
        gf = 512.0 * r["mops"][typ] / 1e9
        tf = gf / r["ms"] if r["ms"] > 0 else 0.0
        busy = r["busy"] / (r["gui"] / nxcc * simds) if r["gui"] > 0 else 0.0
        print("| `%s` | %d | %.4f | %s | %.1f | %.1f | %.3f | %.3f |"
              % (k, r["n"], r["ms"], typ, gf, tf, tf / MFMA_PEAK_TF[typ], busy))


def cmd_counters(a):
    acc = defaultdict(lambda: defaultdict(list))
    with open(a.csv, newline="") as f:
        for r in csv.DictReader(f):
            k = short_name(r["Kernel_Name"])
            if a.match and not any(m in k for m in a.match):
                continue
            acc[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({n for c in acc.values() for n in c})
    print("# %s" % (a.title or "rocprofv3 --pmc counters per dispatch"))
    print()
    if a.command:
        print("Command: `%s`" % a.command)
        print()
    print("Per-dispatch averages; in parentheses, SQ_WAIT_* / SQ_ACTIVE_INST_* as a fraction of SQ_WAVE_CYCLES.")
    print()
    print("| kernel | dispatches | " + " | ".join(names) + " |")
    print("|---|---:|" + "---:|" * len(names))
    for k in sorted(acc):
        c = {n: sum(v) / len(v) for n, v in acc[k].items()}
        wc = c.get("SQ_WAVE_CYCLES", 0.0)
        cells = []
        for n in names:
            v = c.get(n)
            if v is None:
                cells.append("")
            elif wc > 0 and (n.startswith("SQ_WAIT") or n.startswith("SQ_ACTIVE_INST")):
                cells.append("%.4g (%.2f)" % (v, v / wc))
            else:
                cells.append("%.4g" % v)
        print("| `%s` | %d | %s |" % (k, len(next(iter(acc[k].values()))), " | ".join(cells)))


def read_mfma_table(path):
    """{kernel: (dispatches, type, executed GFLOP per dispatch, TFLOP/s, of-peak, busy)} from an `mfma` summary."""
    out = {}
    with open(path) as f:
        for line in f:
            m = re.match(r"\|\s*`([^`]+)`\s*\|(.*)\|\s*$", line)
            if not m:
                continue
            cols = [c.strip() for c in m.group(2).split("|")]
            if len(cols) == 7:
                out[m.group(1)] = (int(cols[0]), cols[2], float(cols[3]), float(cols[4]), float(cols[5]), float(cols[6]))
    return out


def main(argv=None):
    p = argparse.ArgumentParser(prog="profsum.py")
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
    m = sub.add_parser("mfma")
    m.add_argument("csv")
    m.add_argument("--agent", required=True)
    m.add_argument("--title", default="")
    m.add_argument("--command", default="")
    c = sub.add_parser("counters")
    c.add_argument("csv")
    c.add_argument("--match", nargs="*", default=[])
    c.add_argument("--title", default="")
    c.add_argument("--command", default="")
    a = p.parse_args(argv)
    {"stats": cmd_stats, "traffic": cmd_traffic, "mfma": cmd_mfma, "counters": cmd_counters}[a.cmd](a)


if __name__ == "__main__":
    main()
