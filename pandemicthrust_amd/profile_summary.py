"""Summarise rocprofv3 output of the daily loop by the engine's kernel classes.

    python -m pandemicthrust_amd.profile_summary PROFILE_DIR COPY_BYTES > summary.md
    python -m pandemicthrust_amd.profile_summary --kernels STATS_CSV [RUNS] > table.md

PROFILE_DIR holds the csv output of three runs of the same command:
`-o trace` (--kernel-trace --stats), `-o fetch` (--pmc FETCH_SIZE) and
`-o write` (--pmc WRITE_SIZE).  Dispatches are labelled with the classes of
pt_sim_kernel_name by walking each run's dispatch order: kernels before the
first sweep are setup; within a day, the scans between the selection kernel
and the run compaction belong to the selection, and the radix-sort kernels
after the compaction build the day's table (full or subset: the same
launches).  The counters are calibrated with the
snapshot-restore copy kernel (k_copy16), which reads and writes COPY_BYTES
bytes per launch.
"""
import collections
import csv
import os
import sys

SORT = ("k_upsweep", "k_downsweep", "k_scan_reduce", "k_scan_parts", "k_scan_down", "k_offsets")


def short(name):
    n = name.replace("(anonymous namespace)::", "").replace("void ", "").replace("pt::", "")
    return n.split("(")[0].strip()


def label(seq):
    """Class per dispatch for a list of short kernel names in dispatch order."""
    out = []
    section = "setup"
    for n in seq:
        base = n.split("<")[0]
        if base == "k_sweep":
            section = "day"
        if section == "setup":
            cls = "restore" if base in ("k_copy16", "k_fill_u64") else "setup"
        elif base in ("k_copy16", "k_fill_u64"):
            cls = "restore"
        elif base == "k_sweep":
            cls = "sweep"
        elif base == "k_plan":
            cls = "plan"
        elif base == "k_compact":
            cls = "compact"
        elif base in ("k_mark_weekday", "k_mark_weekend", "k_decide"):
            cls = "mark"
        elif base in ("k_gen_weekday", "k_gen_weekend"):
            cls = base[2:]
            section = "full"
        elif base in ("k_select_weekday", "k_select_weekend"):
            cls = "select"
            section = "select"
        elif base == "k_compact_runs":
            cls = "select"
            section = "table"
        elif base in SORT:
            if section == "select":
                cls = "select"
            else:
                cls = {"k_upsweep": "sort_upsweep", "k_downsweep": "sort_downsweep", "k_offsets": "offsets"}.get(
                    base, "sort_scan")
        elif base == "k_contacts":
            cls = "contacts_weekend" if "<true>" in n else "contacts_weekday"
        elif base == "k_apply":
            cls = "apply"
        elif base == "k_end_day":
            cls = "end_day"
        else:
            cls = "other"
        out.append(cls)
    return out


def dispatches(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: int(r.get("Dispatch_Id") or r.get("Correlation_Id")))
    return rows


def counters(path, cname):
    """[(short name, value)] in dispatch order."""
    rows = [r for r in dispatches(path) if r["Counter_Name"] == cname]
    return [(short(r["Kernel_Name"]), float(r["Counter_Value"])) for r in rows]


def main(d, copy_bytes):
    trace = dispatches(os.path.join(d, "trace_kernel_trace.csv"))
    tnames = [short(r["Kernel_Name"]) for r in trace]
    tcls = label(tnames)
    dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3 for r in trace]  # us
    fetch = counters(os.path.join(d, "fetch_counter_collection.csv"), "FETCH_SIZE")
    write = counters(os.path.join(d, "write_counter_collection.csv"), "WRITE_SIZE")

    def calib(vals):
        c = [v for n, v in vals if n == "k_copy16"]
        return copy_bytes / 1024.0 / (sum(c) / len(c))

    kf, kw = calib(fetch), calib(write)
    fcls, wcls = label([n for n, _ in fetch]), label([n for n, _ in write])
    agg = collections.OrderedDict()
    for c, t in zip(tcls, dur):
        a = agg.setdefault(c, {"n": 0, "us": 0.0, "rd": 0.0, "wr": 0.0, "nf": 0, "nw": 0})
        a["n"] += 1
        a["us"] += t
    for c, (_, v) in zip(fcls, fetch):
        if c in agg:
            agg[c]["rd"] += v * 1024 * kf
            agg[c]["nf"] += 1
    for c, (_, v) in zip(wcls, write):
        if c in agg:
            agg[c]["wr"] += v * 1024 * kw
            agg[c]["nw"] += 1
    loop_us = sum(a["us"] for c, a in agg.items() if c not in ("setup", "restore"))
    print("| class | launches | total ms | avg us | % of loop | HBM read MB/launch | HBM write MB/launch |")
    print("|---|---|---|---|---|---|---|")
    for c, a in sorted(agg.items(), key=lambda kv: -kv[1]["us"]):
        rd = a["rd"] / a["nf"] / 1e6 if a["nf"] else float("nan")
        wr = a["wr"] / a["nw"] / 1e6 if a["nw"] else float("nan")
        pct = 100.0 * a["us"] / loop_us if c not in ("setup", "restore") else float("nan")
        print("| %s | %d | %.3f | %.2f | %.1f | %.1f | %.1f |" % (c, a["n"], a["us"] / 1e3, a["us"] / a["n"], pct, rd, wr))
    print()
    print("Loop kernels (classes other than setup and restore): %.1f ms." % (loop_us / 1e3))
    print("Calibration: k_copy16 reads and writes %d bytes per launch; FETCH_SIZE x %.4f, WRITE_SIZE x %.4f "
          "(bytes = KiB x 1024 x factor)." % (copy_bytes, kf, kw))
    s = agg.get("sort_downsweep")
    if s and s["nf"] and s["nw"]:
        per = s["rd"] / s["nf"] + s["wr"] / s["nw"]
        print()
        print("traffic_per_launch sort_downsweep: %.0f bytes over %d launches (avg %.2f us)" % (per, s["n"], s["us"] / s["n"]))


def kernel_table(stats_csv, runs=1):
    """Per-kernel table of a --stats csv (times divided by runs)."""
    rows = list(csv.DictReader(open(stats_csv)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    print("| kernel | calls | ms per run | avg us | % |")
    print("|---|---|---|---|---|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        print("| %s | %s | %.2f | %.1f | %.1f |" % (short(r["Name"]), r["Calls"], float(r["TotalDurationNs"]) / 1e6 / runs,
                                              float(r["AverageNs"]) / 1e3, 100.0 * float(r["TotalDurationNs"]) / tot))
    print()
    print("All kernels: %.1f ms per run." % (tot / 1e6 / runs))


if __name__ == "__main__":
    if sys.argv[1] == "--kernels":
        kernel_table(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 1)
    else:
        main(sys.argv[1], float(sys.argv[2]))
