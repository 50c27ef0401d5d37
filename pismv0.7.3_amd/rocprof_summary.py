"""Turn rocprofv3 CSV output into the short text summaries kept with the
benchmark, and into the HBM traffic per launch of each kernel.

  python pismv0.7.3_amd/rocprof_summary.py stats DIR/run_kernel_stats.csv --cmd "..."
      markdown table of per-kernel call counts and durations
  python pismv0.7.3_amd/rocprof_summary.py traffic --fetch F.csv --write W.csv --copy-bytes N --cmd "..."
      JSON: FETCH_SIZE / WRITE_SIZE per dispatch of every kernel, each scaled by
      the factor that makes the counter of the plain copy kernel (N bytes read,
      N bytes written) read exactly N

FETCH_SIZE and WRITE_SIZE are reported by rocprofv3 in KiB per dispatch.
"""
import argparse
import collections
import csv
import json
import sys

COPY_KERNEL_HINTS = ("copyBuffer", "elementwise_kernel", "copy_kernel")


def short_name(name, width=90):
    name = name.replace("pism_sia::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def kernel_stats(path):
    rows = list(csv.DictReader(open(path)))
    out = []
    for r in rows:
        out.append(dict(name=r["Name"], calls=int(r["Calls"]), total_ns=float(r["TotalDurationNs"]),
                        avg_ns=float(r["AverageNs"]), min_ns=float(r["MinNs"]), max_ns=float(r["MaxNs"]),
                        pct=float(r["Percentage"])))
    out.sort(key=lambda d: -d["total_ns"])
    return out


def stats_markdown(path, cmd):
    lines = ["# rocprofv3 kernel summary", "", "Command:", "", "```", cmd, "```", "",
             "| kernel | calls | avg ms | min ms | max ms | % of GPU time |",
             "|---|---:|---:|---:|---:|---:|"]
    for d in kernel_stats(path):
        lines.append("| `%s` | %d | %.4f | %.4f | %.4f | %.2f |" % (
            short_name(d["name"]), d["calls"], d["avg_ns"] / 1e6, d["min_ns"] / 1e6, d["max_ns"] / 1e6, d["pct"]))
    return "\n".join(lines) + "\n"


def counter_per_kernel(path, counter):
    """kernel name -> list of counter values (bytes) per dispatch."""
    vals = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        vals[r["Kernel_Name"]].append(float(r["Counter_Value"]) * 1024.0)
    return vals


def copy_factor(vals, copy_bytes):
    """copy_bytes / counter of the largest plain copy dispatch."""
    best = 0.0
    for name, v in vals.items():
        if any(h in name for h in COPY_KERNEL_HINTS):
            best = max(best, max(v))
    if best <= 0.0:
        raise SystemExit("no copy kernel dispatch found for calibration")
    return copy_bytes / best


def traffic(fetch_csv, write_csv, copy_bytes):
    fv = counter_per_kernel(fetch_csv, "FETCH_SIZE")
    wv = counter_per_kernel(write_csv, "WRITE_SIZE")
    ff = copy_factor(fv, copy_bytes)
    fw = copy_factor(wv, copy_bytes)
    out = {"copy_bytes": copy_bytes, "fetch_scale": ff, "write_scale": fw, "kernels": {}}
    for name in sorted(set(fv) | set(wv)):
        if any(h in name for h in COPY_KERNEL_HINTS):
            continue
        f = fv.get(name, [0.0])
        w = wv.get(name, [0.0])
        fb = ff * sum(f) / len(f)
        wb = fw * sum(w) / len(w)
        out["kernels"][short_name(name, 200)] = {"fetch_bytes": fb, "write_bytes": wb,
                                                 "bytes_per_launch": fb + wb, "dispatches": len(f)}
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--cmd", required=True)
    b = sub.add_parser("traffic")
    b.add_argument("--fetch", required=True)
    b.add_argument("--write", required=True)
    b.add_argument("--copy-bytes", type=float, required=True)
    b.add_argument("--cmd", required=True)
    args = ap.parse_args(argv)
    if args.what == "stats":
        sys.stdout.write(stats_markdown(args.csv, args.cmd))
    else:
        t = traffic(args.fetch, args.write, args.copy_bytes)
        t["command"] = args.cmd
        json.dump(t, sys.stdout, indent=1)
        sys.stdout.write("\n")


if __name__ == "__main__":
    main()
