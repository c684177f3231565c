"""Summaries of rocprofv3 CSV output for this library's kernels.

    python -m pydyn_amd.profiling stats DIR
        the kernel statistics of a `rocprofv3 --kernel-trace --stats --output-format csv` run
    python -m pydyn_amd.profiling traffic FETCH_DIR WRITE_DIR COPY_BYTES [--json OUT]
        HBM bytes per launch of every kernel from a `--pmc FETCH_SIZE` run and a `--pmc
        WRITE_SIZE` run of the same command, each scaled by what the counter reads for a plain
        device copy of COPY_BYTES bytes in that run (the largest __amd_rocclr_copyBuffer
        dispatch): bytes = FETCH_KiB * 1024 * (COPY_BYTES / copy FETCH bytes) + likewise WRITE.
    python -m pydyn_amd.profiling counters DIR [DIR ...] --kernel REGEX
        the mean per dispatch of every counter of the `--pmc` runs in the DIRs for the kernels
        whose name matches REGEX, and the ratios the SQ counters give (fractions of the waves'
        cycles spent issuing VALU or LDS instructions or waiting, FP64 instructions per wave).
"""

import argparse
import collections
import csv
import glob
import json
import os
import re
import sys


def _short(kernel):
    """The library's own name of a kernel (tm5_profile_*): the column kernel's template flag
    dropped, its adjoint named as such."""
    short = kernel.split("(")[0].replace("void ", "").replace("tm5::", "")
    m = re.fullmatch(r"k_column<(\d+), (true|false)>", short)
    if m:
        return f"k_column<{m.group(1)}>" if m.group(2) == "false" else "k_column_adjoint"
    return short


def _csv(d, suffix):
    found = sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))
    if not found:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    with open(found[0], newline="") as f:
        return list(csv.DictReader(f))


def stats(d):
    """[(kernel, calls, total_ns, average_ns, percent)] from kernel_stats.csv, largest first."""
    rows = []
    for r in _csv(d, "kernel_stats.csv"):
        rows.append((r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                     float(r["Percentage"])))  # fmt: skip
    return sorted(rows, key=lambda x: -x[2])


def medians(d):
    """{kernel: median duration ns} from kernel_trace.csv, when the run kept its trace: the launch
    a kernel makes when it works (a queued run's launches after a missed plan skip in about a
    microsecond and pull the average down)."""
    durs = collections.defaultdict(list)
    try:
        rows = _csv(d, "kernel_trace.csv")
    except FileNotFoundError:
        return {}
    for r in rows:
        durs[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    out = {}
    for k, v in durs.items():
        v.sort()
        out[k] = float(v[len(v) // 2])
    return out


def counter(d, name):
    """{kernel: [dispatches, mean value]} of one counter from counter_collection.csv."""
    acc = collections.defaultdict(list)
    for r in _csv(d, "counter_collection.csv"):
        if r["Counter_Name"] == name:
            acc[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return {k: [len(v), sum(v) / len(v), max(v)] for k, v in acc.items()}


def traffic(fetch_dir, write_dir, copy_bytes):
    fetch, write = counter(fetch_dir, "FETCH_SIZE"), counter(write_dir, "WRITE_SIZE")
    copy = [k for k in fetch if "copyBuffer" in k]
    if not copy:
        raise ValueError("no copy kernel in the FETCH_SIZE run")
    kf = copy_bytes / (max(fetch[k][2] for k in copy) * 1024.0)
    copy_w = [k for k in write if "copyBuffer" in k]
    kw = copy_bytes / (max(write[k][2] for k in copy_w) * 1024.0)
    out = {"_calibration": {"copy_bytes": copy_bytes, "fetch_factor": round(kf, 4), "write_factor": round(kw, 4)}}
    for k in fetch:
        if k in write:
            b = fetch[k][1] * 1024.0 * kf + write[k][1] * 1024.0 * kw
            short = _short(k)
            out[short] = {"bytes_per_launch": round(b), "fetch_kib": round(fetch[k][1], 1),
                          "write_kib": round(write[k][1], 1), "dispatches": fetch[k][0]}  # fmt: skip
    return out


def counters(dirs, kernel):
    """{counter: (dispatches, mean per dispatch)} over the --pmc runs in dirs, for the kernels whose
    name matches the regex kernel."""
    pat = re.compile(kernel)
    acc = collections.defaultdict(list)
    for d in dirs:
        for r in _csv(d, "counter_collection.csv"):
            if pat.search(r["Kernel_Name"]):
                acc[r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: (len(v), sum(v) / len(v)) for k, v in sorted(acc.items())}


RATIOS = [  # (label, numerator, denominator)
    ("VALU issue / wave cycles", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES"),
    ("LDS issue / wave cycles", "SQ_ACTIVE_INST_LDS", "SQ_WAVE_CYCLES"),
    ("any issue / wave cycles", "SQ_ACTIVE_INST_ANY", "SQ_WAVE_CYCLES"),
    ("waiting (waitcnt) / wave cycles", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"),
    ("waiting to issue / wave cycles", "SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES"),
    ("waiting to issue LDS / wave cycles", "SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES"),
    ("VALU instructions per wave", "SQ_INSTS_VALU", "SQ_WAVES"),
    ("FP64 add per wave", "SQ_INSTS_VALU_ADD_F64", "SQ_WAVES"),
    ("FP64 mul per wave", "SQ_INSTS_VALU_MUL_F64", "SQ_WAVES"),
    ("FP64 fma per wave", "SQ_INSTS_VALU_FMA_F64", "SQ_WAVES"),
    ("FP64 transcendental per wave", "SQ_INSTS_VALU_TRANS_F64", "SQ_WAVES"),
    ("LDS instructions per wave", "SQ_INSTS_LDS", "SQ_WAVES"),
    ("LDS bank conflict cycles / LDS issue", "SQ_LDS_BANK_CONFLICT", "SQ_ACTIVE_INST_LDS"),
    ("busy / GPU active cycles", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE"),
]


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = p.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("copy_bytes", type=int)
    t.add_argument("--json")
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--kernel", required=True)
    a = p.parse_args(argv)
    if a.cmd == "counters":
        got = counters(a.dirs, a.kernel)
        print(f"{'counter':32s} {'dispatches':>10s} {'mean per dispatch':>20s}")
        for k, (n, mean) in got.items():
            print(f"{k:32s} {n:10d} {mean:20.1f}")
        for label, num, den in RATIOS:
            if num in got and den in got and got[den][1] > 0:
                print(f"{label:40s} {got[num][1] / got[den][1]:10.3f}")
        return 0
    if a.cmd == "stats":
        med = medians(a.dir)
        print(f"{'kernel':64s} {'calls':>6s} {'avg_us':>9s} {'median_us':>9s} {'total_us':>10s} {'pct':>6s}")
        for name, calls, tot, avg, pct in stats(a.dir):
            short = name.split("(")[0].replace("void ", "")
            mu = f"{med[name] / 1e3:9.2f}" if name in med else f"{'-':>9s}"
            print(f"{short[:64]:64s} {calls:6d} {avg / 1e3:9.2f} {mu} {tot / 1e3:10.1f} {pct:6.2f}")
    else:
        out = traffic(a.fetch_dir, a.write_dir, a.copy_bytes)
        if a.json:
            with open(a.json, "w") as f:
                f.write(json.dumps(out, indent=1, sort_keys=True) + "\n")
        print(f"calibration: {out['_calibration']}")
        print(f"{'kernel':36s} {'MB/launch':>10s} {'fetch KiB':>10s} {'write KiB':>10s} {'dispatches':>10s}")
        for k in sorted(out):
            if k.startswith("_") or "rocclr" in k or "at::" in k:
                continue
            v = out[k]
            print(f"{k[:36]:36s} {v['bytes_per_launch'] / 1e6:10.2f} {v['fetch_kib']:10.1f} {v['write_kib']:10.1f} "
                  f"{v['dispatches']:10d}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
