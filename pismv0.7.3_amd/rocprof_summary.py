"""Turn rocprofv3 CSV output into the short text summaries kept with the
benchmark, and into the HBM traffic per launch of each kernel.

  python pismv0.7.3_amd/rocprof_summary.py stats DIR/run_kernel_stats.csv --cmd "..."
      markdown table of per-kernel call counts and durations
  python pismv0.7.3_amd/rocprof_summary.py traffic --fetch F.csv --write W.csv --copy-bytes N --cmd "..."
      JSON: FETCH_SIZE / WRITE_SIZE per dispatch of every kernel, each scaled by
      the factor that makes the counter of the plain copy kernel (N bytes read,
      N bytes written) read exactly N

  python pismv0.7.3_amd/rocprof_summary.py valu P1.csv [P2.csv ...] --cmd "..."
      markdown table of VALU issue utilisation, FP64 instruction mix and rate
      and lane utilisation per kernel (longest dispatch of each), from the SQ
      counters SQ_INSTS_VALU, SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64,
      SQ_INSTS_VALU_FLOPS_FP64, SQ_ACTIVE_INST_VALU, SQ_THREAD_CYCLES_VALU and
      GRBM_GUI_ACTIVE (any split of them over several passes)

FETCH_SIZE and WRITE_SIZE are reported by rocprofv3 in KiB per dispatch.

VALU figures on MI355X (256 CUs, 4 SIMD16 units each, 8 XCDs): a wave64 VALU
instruction occupies its SIMD for 4 cycles, so VALU issue utilisation =
4 SQ_INSTS_VALU / (1024 x kernel cycles), kernel cycles = GRBM_GUI_ACTIVE / 8
(the counter sums the 8 XCDs). SQ_INSTS_VALU_FLOPS_FP64 counts FLOPs per
wave instruction (FMA = 2); with full wavefronts the FP64 rate is 64 x that
over the dispatch time, against the 78.6 TFLOP/s FP64 vector peak. Lane
utilisation = SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU).
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


FP64_VECTOR_PEAK_TFLOPS = 78.6
N_SIMD, N_XCD = 1024, 8


def valu_counters(paths):
    """{kernel: {counter: value}} of the longest dispatch of each kernel, merged over passes."""
    out = {}
    for path in paths:
        per = collections.defaultdict(lambda: collections.defaultdict(dict))
        dur = collections.defaultdict(dict)
        for r in csv.DictReader(open(path)):
            d = int(r["Dispatch_Id"])
            per[r["Kernel_Name"]][d][r["Counter_Name"]] = float(r["Counter_Value"])
            dur[r["Kernel_Name"]][d] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
        for name, disp in per.items():
            longest = max(dur[name], key=lambda k: dur[name][k])
            o = out.setdefault(name, {})
            o.update(disp[longest])
            o.setdefault("seconds", dur[name][longest])
    return out


def valu_markdown(paths, cmd, min_seconds=1e-4):
    c = valu_counters(paths)
    lines = ["# VALU and FP64 counters per kernel (longest dispatch)", "", "Commands:", "", "```", cmd, "```", "",
             "| kernel | ms | VALU instr | VALU issue util | FP64 instr share | FP64 TFLOP/s | % of FP64 peak | lane util |",
             "|---|---:|---:|---:|---:|---:|---:|---:|"]
    rows = []
    for name, v in c.items():
        if v.get("seconds", 0) < min_seconds or "SQ_INSTS_VALU" not in v or "GRBM_GUI_ACTIVE" not in v:
            continue
        cyc = v["GRBM_GUI_ACTIVE"] / N_XCD
        issue = 4.0 * v["SQ_INSTS_VALU"] / (N_SIMD * cyc) if cyc > 0 else 0.0
        f64 = sum(v.get("SQ_INSTS_VALU_%s_F64" % k, 0.0) for k in ("ADD", "MUL", "FMA", "TRANS"))
        share = f64 / v["SQ_INSTS_VALU"] if v["SQ_INSTS_VALU"] else 0.0
        tf = 64.0 * v.get("SQ_INSTS_VALU_FLOPS_FP64", 0.0) / v["seconds"] / 1e12
        lane = (v["SQ_THREAD_CYCLES_VALU"] / (64.0 * v["SQ_ACTIVE_INST_VALU"])
                if v.get("SQ_ACTIVE_INST_VALU") and "SQ_THREAD_CYCLES_VALU" in v else float("nan"))
        rows.append((v["seconds"], "| `%s` | %.3f | %.3g | %.1f%% | %.1f%% | %.1f | %.1f%% | %.1f%% |" % (
            short_name(name), 1e3 * v["seconds"], v["SQ_INSTS_VALU"], 100 * issue, 100 * share, tf,
            100 * tf / FP64_VECTOR_PEAK_TFLOPS, 100 * lane)))
    rows.sort(key=lambda r: -r[0])
    lines += [r for _, r in rows]
    return "\n".join(lines) + "\n"


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
    v = sub.add_parser("valu")
    v.add_argument("csv", nargs="+")
    v.add_argument("--cmd", required=True)
    args = ap.parse_args(argv)
    if args.what == "stats":
        sys.stdout.write(stats_markdown(args.csv, args.cmd))
    elif args.what == "valu":
        sys.stdout.write(valu_markdown(args.csv, args.cmd))
    else:
        t = traffic(args.fetch, args.write, args.copy_bytes)
        t["command"] = args.cmd
        json.dump(t, sys.stdout, indent=1)
        sys.stdout.write("\n")


if __name__ == "__main__":
    main()
