"""Turns rocprofv3 CSV output into the short summaries kept in profiles/.

  python3 profile_summary.py RAW DST                 (profile_bench.sh: volume-render frame)
  python3 profile_summary.py --stats RAW OUT TITLE COMMAND   (profile_workloads.sh)
"""
import csv
import glob
import os
import sys
from collections import defaultdict

RAW = DST = None
CMD = ("python3 bench.py --steps 5 --warmup 1 --no-cpu --no-extra --no-batched --no-served --trace-config 16,256,32 --volume-layout 4 --copy-calibration "
       "(volume render 1920x1080, 512^3 noise, iso 0.5 + shadows)")


def rows(d, pat="*counter_collection.csv"):
    f = glob.glob(os.path.join(RAW, d, "**", pat), recursive=True)
    return list(csv.DictReader(open(f[0]))) if f else []


def main():
    out = []
    st = glob.glob(os.path.join(RAW, "ktrace", "**", "*kernel_stats.csv"), recursive=True)
    lines = ["# Kernel trace, volume-render frame", "",
             "Made by `bash profile_bench.sh` (repository root), which runs",
             "`rocprofv3 --kernel-trace --stats --output-format csv -- " + CMD + "`", "",
             "| Kernel | Calls | Total ms | Average us | % |", "|---|---:|---:|---:|---:|"]
    if st:
        for r in sorted(csv.DictReader(open(st[0])), key=lambda r: -float(r["TotalDurationNs"]))[:18]:
            lines.append("| `%s` | %s | %.3f | %.1f | %.1f |" % (r["Name"][:70], r["Calls"],
                                                                float(r["TotalDurationNs"]) / 1e6,
                                                                float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    open(os.path.join(DST, "trace_volume_render_kernels.md"), "w").write("\n".join(lines) + "\n")
    f, w = rows("fetch"), rows("write")
    cf = max(float(r["Counter_Value"]) for r in f if "copyBuffer" in r["Kernel_Name"])
    cw = max(float(r["Counter_Value"]) for r in w if "copyBuffer" in r["Kernel_Name"])
    sf, sw = 256 * 1024 / cf, 256 * 1024 / cw
    kf = [float(r["Counter_Value"]) for r in f if "k_trace<1" in r["Kernel_Name"]]
    kw = [float(r["Counter_Value"]) for r in w if "k_trace<1" in r["Kernel_Name"]]
    tb = [(a * sf + b * sw) * 1024 for a, b in zip(kf, kw)]
    sq = defaultdict(list)
    for r in rows("sq"):
        if "k_trace<1" in r["Kernel_Name"]:
            sq[r["Counter_Name"]].append(float(r["Counter_Value"]))
    tcc = defaultdict(list)
    for r in rows("tcc"):
        if "k_trace<1" in r["Kernel_Name"]:
            tcc[r["Counter_Name"]].append(float(r["Counter_Value"]))
    avg = lambda v: sum(v) / len(v) if v else float("nan")  # noqa: E731
    util = avg(sq["SQ_THREAD_CYCLES_VALU"]) / (64 * avg(sq["SQ_ACTIVE_INST_VALU"]))
    hit = avg(tcc["TCC_HIT_sum"]) / (avg(tcc["TCC_HIT_sum"]) + avg(tcc["TCC_MISS_sum"]))
    lines = [
        "# HBM traffic of the trace kernel (volume-render frame)", "",
        "Made by `bash profile_bench.sh` (repository root); one counter group per run:",
        "`rocprofv3 --pmc FETCH_SIZE --output-format csv -- " + CMD + "`",
        "`rocprofv3 --pmc WRITE_SIZE --output-format csv -- " + CMD + "`", "",
        "Calibration: one 256 MiB device-to-device copy in the same run reads FETCH_SIZE %.0f KiB and "
        "WRITE_SIZE %.0f KiB, so FETCH_SIZE is scaled by %.4f and WRITE_SIZE by %.4f." % (cf, cw, sf, sw), "",
        "trace launches measured = %d (primary and shadow passes)" % len(kf),
        "fetch_bytes_per_launch = %.0f" % (avg(kf) * sf * 1024),
        "write_bytes_per_launch = %.0f" % (avg(kw) * sw * 1024),
        "traffic_bytes_per_launch = %.0f" % avg(tb), "",
        "SQ counters (same kernel, own run): VALU lane utilisation %.3f "
        "(SQ_THREAD_CYCLES_VALU / 64 SQ_ACTIVE_INST_VALU), VALU instructions per wave %.0f, "
        "waves per launch %.0f" % (util, avg(sq["SQ_INSTS_VALU"]) / max(1, avg(sq["SQ_WAVES"])), avg(sq["SQ_WAVES"])),
        "L2 (TCC) hit rate %.3f" % hit,
    ]
    open(os.path.join(DST, "trace_volume_render_traffic.txt"), "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


def stats_table(raw, out, title, cmd, top=24):
    st = glob.glob(os.path.join(raw, "**", "*kernel_stats.csv"), recursive=True)
    lines = ["# " + title, "", "Made by `bash profile_workloads.sh` (repository root), which runs",
             "`rocprofv3 --kernel-trace --stats --output-format csv -- " + cmd + "`", "",
             "| Kernel | Calls | Total ms | Average us | % |", "|---|---:|---:|---:|---:|"]
    for r in sorted(csv.DictReader(open(st[0])), key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        lines.append("| `%s` | %s | %.3f | %.1f | %.1f |" % (r["Name"][:80], r["Calls"],
                                                            float(r["TotalDurationNs"]) / 1e6,
                                                            float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    if sys.argv[1] == "--stats":
        stats_table(*sys.argv[2:6])
    else:
        RAW, DST = sys.argv[1], sys.argv[2]
        main()
