"""Turn rocprofv3 --pmc CSV output into per-kernel HBM traffic and instruction-mix figures, or into a table of
any counters per kernel.

    python -m idgx.pmc_summary --cal-fetch DIR --cal-write DIR --fetch DIR --write DIR [--sq DIR] \
        --cal-mib 1024 --out profiles/hbm_traffic_ska_low
    python -m idgx.pmc_summary --counters DIR [DIR ...] --out profiles/pipe_utilisation_ska_low [--title ...]
    python -m idgx.pmc_summary --stats DIR --out profiles/gridder_degridder_ska_low [--title ...] [--notes ...]

FETCH_SIZE / WRITE_SIZE are read in KiB per dispatch.  The calibration runs are a plain device-to-device copy of
--cal-mib MiB (idgx.calibrate); scale factor = true bytes / counter bytes of that copy, applied to the kernels'
counters.  Writes <out>.md (table, commands) and <out>.json (bytes per launch by kernel), both plain text.
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys

KERNELS = ["gridder_plain_queue_tiles", "gridder_plain_queue", "degridder_plain_queue", "gridder_plain",
           "degridder_plain", "gridder_mfma_queue_tiles", "gridder_mfma_queue", "degridder_mfma_queue",
           "degrid_prepare_kernel", "gridder_mfma", "degridder_mfma", "subgrid_fft_kernel", "adder_kernel",
           "splitter_kernel", "fft_col_kernel", "fft_row_kernel", "wstack_row_kernel", "copy_box_kernel",
           "clear_band_kernel",
           "checkerboard_kernel", "wstack_to_image_kernel", "wstack_from_image_kernel", "reset_counters_kernel"]


def short_name(name):
    """Kernel family plus template arguments ("gridder_mfma_queue<1, true, 8, 256, false>"), so that every
    instantiation (precision, tile shape) keeps its own row; "copy" for the calibration copy."""
    for k in KERNELS:
        i = name.find(k)
        if i < 0 or (k.startswith("gridder") and name[i - 2:i] == "de"):
            continue
        if name[i + len(k):i + len(k) + 1] not in ("<", "(", ""):
            continue
        j = i + len(k)
        if name[j:j + 1] == "<":
            depth = 0
            for q in range(j, len(name)):
                depth += {"<": 1, ">": -1}.get(name[q], 0)
                if depth == 0:
                    return name[i:q + 1]
        return k
    if "copyBuffer" in name or "copy" in name.lower():
        return "copy"
    return None


def load(d):
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for p in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for r in csv.DictReader(open(p)):
            k = short_name(r["Kernel_Name"])
            if k:
                out[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return out


def mean(v):
    return sum(v) / len(v) if v else float("nan")


def counter_table(dirs, out, title, command):
    """Mean per dispatch of every counter in `dirs` (one rocprofv3 --pmc run each), one row per kernel."""
    per = collections.defaultdict(dict)
    names = []
    for d in dirs:
        for k, counters in load(d).items():
            for c, v in counters.items():
                per[k][c] = v
                if c not in names:
                    names.append(c)
    lines = [f"# {title}", ""]
    if command:
        lines += ["Commands:", "", "```", command, "```", ""]
    lines += ["| kernel | launches | " + " | ".join(names) + " |", "|---|---|" + "---|" * len(names)]
    res = {}
    for k in sorted(per):
        if k == "copy":
            continue
        n = max(len(v) for v in per[k].values())
        res[k] = {c: mean(v) for c, v in per[k].items()}
        lines.append(f"| `{k}` | {n} | " + " | ".join(f"{mean(per[k].get(c, [])):.4g}" for c in names) + " |")
    with open(out + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(out + ".json", "w") as f:
        json.dump(res, f, indent=1)
    print("\n".join(lines))
    return 0


def stats_table(d, out, title, command, notes=""):
    """rocprofv3 --kernel-trace --stats output (the kernel_stats.csv of run directory d): calls, mean / min / max
    duration and share per kernel, this library's kernels by family and template arguments, others by name."""
    paths = glob.glob(os.path.join(d, "**", "*kernel_stats.csv"), recursive=True)
    if not paths:
        raise SystemExit(f"no kernel_stats.csv under {d}")
    rows = []
    for r in csv.DictReader(open(paths[0])):
        name = short_name(r["Name"]) or r["Name"].split("(")[0].replace("void ", "")[:60]
        rows.append((name, int(r["Calls"]), float(r["AverageNs"]) / 1e6, float(r["MinNs"]) / 1e6,
                     float(r["MaxNs"]) / 1e6, float(r["Percentage"])))
    rows.sort(key=lambda x: -x[5])
    lines = [f"# {title}", ""]
    if notes:
        lines += [notes, ""]
    if command:
        lines += ["Command (run from /tmp with TMPDIR=/tmp, repository at $R):", "", "```", command, "```", ""]
    lines += ["| kernel | calls | average (ms) | min (ms) | max (ms) | share |", "|---|---|---|---|---|---|"]
    for name, calls, avg, mn, mx, pct in rows:
        if pct < 0.01:
            continue
        lines.append(f"| `{name}` | {calls} | {avg:.3f} | {mn:.3f} | {mx:.3f} | {pct:.2f} % |")
    with open(out + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", help="tabulate a rocprofv3 --kernel-trace --stats run directory")
    ap.add_argument("--notes", default="", help="text placed under the title")
    for k in ("cal-fetch", "cal-write", "fetch", "write"):
        ap.add_argument("--" + k)
    ap.add_argument("--sq")
    ap.add_argument("--counters", nargs="+", help="tabulate every counter of these runs per kernel")
    ap.add_argument("--title", default="Counters per kernel (rocprofv3 --pmc)")
    ap.add_argument("--cal-mib", type=float, default=1024)
    ap.add_argument("--out", required=True)
    ap.add_argument("--command", default="")
    a = ap.parse_args()
    if a.stats:
        return stats_table(a.stats, a.out, a.title, a.command, a.notes)
    if a.counters:
        return counter_table(a.counters, a.out, a.title, a.command)
    if not (a.cal_fetch and a.cal_write and a.fetch and a.write):
        ap.error("--cal-fetch, --cal-write, --fetch and --write are required without --counters")
    true_bytes = a.cal_mib * 1024 * 1024
    f_fetch = true_bytes / (mean(load(a.cal_fetch)["copy"]["FETCH_SIZE"]) * 1024)
    f_write = true_bytes / (mean(load(a.cal_write)["copy"]["WRITE_SIZE"]) * 1024)
    fetch, write = load(a.fetch), load(a.write)
    sq = load(a.sq) if a.sq else {}
    res = {"calibration": {"copy_mib": a.cal_mib, "fetch_factor": f_fetch, "write_factor": f_write}, "kernels": {}}
    lines = ["# HBM traffic and instruction mix per kernel (rocprofv3 --pmc)", "",
             f"Calibration: device-to-device copy of {a.cal_mib:.0f} MiB (idgx.calibrate); FETCH_SIZE factor "
             f"{f_fetch:.3f}, WRITE_SIZE factor {f_write:.3f} (true bytes / counter bytes).", ""]
    if a.command:
        lines += ["Commands:", "", "```", a.command, "```", ""]
    if sq:
        lines += ["| kernel | launches | HBM read GB/launch | HBM write GB/launch | VALU instr | TRANS_F32 | FMA_F32 | "
                  "MFMA MOPs F16 | MFMA MOPs BF16 |",
                  "|---|---|---|---|---|---|---|---|---|"]
    else:  # traffic only
        lines += ["| kernel | launches | HBM read GB/launch | HBM write GB/launch |", "|---|---|---|---|"]
    for k in sorted(set(fetch) | set(write)):
        if k == "copy":
            continue
        rd = mean(fetch[k]["FETCH_SIZE"]) * 1024 * f_fetch if k in fetch else float("nan")
        wr = mean(write[k]["WRITE_SIZE"]) * 1024 * f_write if k in write else float("nan")
        s = sq.get(k, {})
        res["kernels"][k] = {"read_bytes": rd, "write_bytes": wr, "traffic_bytes": rd + wr,
                             **{c: mean(v) for c, v in s.items()}}
        n = len(fetch[k]["FETCH_SIZE"]) if k in fetch else len(write[k]["WRITE_SIZE"])
        res["kernels"][k]["launches"] = n
        if not sq:
            lines.append(f"| `{k}` | {n} | {rd / 1e9:.2f} | {wr / 1e9:.2f} |")
            continue
        lines.append(f"| `{k}` | {n} | {rd / 1e9:.2f} | {wr / 1e9:.2f} | {mean(s.get('SQ_INSTS_VALU', [])):.3e} | "
                     f"{mean(s.get('SQ_INSTS_VALU_TRANS_F32', [])):.3e} | {mean(s.get('SQ_INSTS_VALU_FMA_F32', [])):.3e} | "
                     f"{mean(s.get('SQ_INSTS_VALU_MFMA_MOPS_F16', [])):.3e} | "
                     f"{mean(s.get('SQ_INSTS_VALU_MFMA_MOPS_BF16', [])):.3e} |")
    with open(a.out + ".md", "w") as f:
        f.write("\n".join(lines) + "\n")
    with open(a.out + ".json", "w") as f:
        json.dump(res, f, indent=1)
    print("\n".join(lines))
    return 0


if __name__ == "__main__":
    sys.exit(main())
