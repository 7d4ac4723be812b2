"""Turns rocprofv3 CSV output into the short summaries kept in profiles/.

  python3 profile_summary.py RAW DST                 (profile_bench.sh: volume-render frame)
  python3 profile_summary.py --stats RAW OUT TITLE COMMAND   (profile_workloads.sh)
  python3 profile_summary.py --counters RAW OUT TITLE COMMAND (profile_workloads.sh: geometry kernels)
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


def counters_table(raw, out, title, cmd):
    """Per geometry trace kernel (k_trace<0, true, ...>): SQ, FETCH_SIZE and
    TCC counters of its launches (three runs of the same command), FETCH_SIZE
    scaled by what a 256 MiB device copy in the same run reads."""
    def load(d):
        f = glob.glob(os.path.join(raw, d, "**", "*counter_collection.csv"), recursive=True)
        return list(csv.DictReader(open(f[0]))) if f else []
    fetch = load("mesh_fetch")
    cf = max((float(r["Counter_Value"]) for r in fetch if "copyBuffer" in r["Kernel_Name"]), default=float("nan"))
    scale = 256 * 1024 / cf
    per = defaultdict(lambda: defaultdict(list))
    for d in ("mesh_sq", "mesh_fetch", "mesh_tcc"):
        for r in load(d):
            k = r["Kernel_Name"]
            if "k_trace<0, true" in k:
                per[k.split("(")[0]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    avg = lambda v: sum(v) / len(v) if v else float("nan")  # noqa: E731
    lines = ["# " + title, "", "Made by `bash profile_workloads.sh` (repository root); one counter group per run:"]
    for grp in ("SQ_WAVES SQ_INSTS_VALU SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_INSTS_VMEM_RD SQ_WAVE_CYCLES "
                "SQ_BUSY_CYCLES SQ_INSTS_SALU", "FETCH_SIZE", "TCC_HIT_sum TCC_MISS_sum"):
        lines.append("`rocprofv3 --pmc %s --output-format csv -- %s`" % (grp, cmd))
    lines += ["", "FETCH_SIZE scaled by %.4f (a 256 MiB device-to-device copy in the same run reads %.0f KiB)."
              % (scale, cf), "",
              "| Kernel | Launches | VALU lane utilisation | VALU instr / wave | VMEM reads / wave | Waves / launch "
              "| L2 hit rate | HBM fetch MB / launch |", "|---|---:|---:|---:|---:|---:|---:|---:|"]
    for k in sorted(per):
        c = per[k]
        waves = avg(c["SQ_WAVES"])
        util = avg(c["SQ_THREAD_CYCLES_VALU"]) / (64 * avg(c["SQ_ACTIVE_INST_VALU"]))
        hit = avg(c["TCC_HIT_sum"]) / (avg(c["TCC_HIT_sum"]) + avg(c["TCC_MISS_sum"]))
        lines.append("| `%s` | %d | %.3f | %.0f | %.0f | %.0f | %.3f | %.1f |" % (
            k, len(c["SQ_WAVES"]), util, avg(c["SQ_INSTS_VALU"]) / waves, avg(c["SQ_INSTS_VMEM_RD"]) / waves,
            waves, hit, avg(c["FETCH_SIZE"]) * scale * 1024 / 1e6))
    open(out, "w").write("\n".join(lines) + "\n")
    print("\n".join(lines))


if __name__ == "__main__":
    if sys.argv[1] == "--stats":
        stats_table(*sys.argv[2:6])
    elif sys.argv[1] == "--counters":
        counters_table(*sys.argv[2:6])
    else:
        RAW, DST = sys.argv[1], sys.argv[2]
        main()
