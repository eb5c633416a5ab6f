"""Summarise rocprofv3 CSV output for the chemistry kernels.

    python -m nrg_amd.rocprof_summary stats   <kernel_stats.csv>
    python -m nrg_amd.rocprof_summary trace   <kernel_trace.csv> --last N
    python -m nrg_amd.rocprof_summary counter <counter_collection.csv> [--calib-mib M] [--last N]

``stats`` prints calls / average / total duration per kernel.  ``trace``
prints, per kernel, the mean duration of its last N dispatches (the timed
steps, after warm-up) from the per-dispatch trace.  ``counter``
prints, per kernel, the average per-dispatch value of each collected counter
(FETCH_SIZE and WRITE_SIZE are in KiB per dispatch), over its last N dispatches
with ``--last N`` (the timed steps), else over all of them.  With
SQ_THREAD_CYCLES_VALU and SQ_ACTIVE_INST_VALU both collected, a line
    LANES <tab> kernel <tab> share of the 64 lanes busy per VALU cycle
follows for each kernel.  With ``--calib-mib`` the
largest single device-copy dispatch (the known-size copy bench.py runs with
--copy-calibration-mib) is taken as the reference: its expected bytes divided
by its measured bytes gives the factor applied to the other kernels.  Each
calibrated byte counter also gets a line
    TRAFFIC <tab> kernel <tab> counter <tab> bytes per dispatch
which bench.py reads to report the HBM traffic of the chemistry kernels.
"""
import csv
import re
import sys
from collections import defaultdict


def _short(name, width=70):
    return name if len(name) <= width else name[: width - 3] + "..."


def stats(path):
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("%-72s %6s %14s %14s %7s" % ("kernel", "calls", "avg_ms", "total_ms", "pct"))
    for r in rows[:15]:
        print("%-72s %6s %14.3f %14.3f %7.2f" % (_short(r["Name"]), r["Calls"],
                                                float(r["AverageNs"]) / 1e6,
                                                float(r["TotalDurationNs"]) / 1e6,
                                                float(r["Percentage"])))


def trace(path, last):
    rows = list(csv.DictReader(open(path)))
    name_key = "Kernel_Name" if "Kernel_Name" in rows[0] else "Kernel-Name"
    per = defaultdict(list)
    for r in rows:
        b = float(r.get("Start_Timestamp") or r.get("Start-Timestamp"))
        e = float(r.get("End_Timestamp") or r.get("End-Timestamp"))
        per[r[name_key]].append((b, e - b))
    print("%-72s %6s %14s   (mean of the last %d dispatches)" % ("kernel", "disp", "avg_ms", last))
    for k in sorted(per, key=lambda k: -sum(d for _, d in per[k])):
        ds = [d for _, d in sorted(per[k])]
        tail = ds[-last:]
        print("%-72s %6d %14.3f" % (_short(k), len(ds), sum(tail) / len(tail) / 1e6))


def counters(path, calib_mib=None, last=None):
    rows = list(csv.DictReader(open(path)))
    name_key = "Kernel_Name" if "Kernel_Name" in rows[0] else "Kernel-Name"
    per_disp = defaultdict(float)          # (kernel, counter, dispatch) -> value
    for r in rows:
        k = r[name_key]
        cname = r.get("Counter_Name") or r.get("Counter-Name")
        did = r.get("Dispatch_Id") or r.get("Dispatch-Id")
        val = float(r.get("Counter_Value") or r.get("Counter-Value"))
        per_disp[(k, cname, did)] += val
    # each kernel's dispatches in launch order; with --last only the last N count
    kd = defaultdict(set)
    for (k, _, did) in per_disp:
        kd[k].add(int(did))
    keep = {k: set(sorted(v)[-last:] if last else v) for k, v in kd.items()}
    acc = defaultdict(lambda: defaultdict(float))
    disp = defaultdict(set)
    for (k, cname, did), val in per_disp.items():
        if int(did) in keep[k]:
            acc[k][cname] += val
            disp[k].add(did)
    factor = {}
    if calib_mib:
        # bench.py's known-size device-to-device copy runs as the runtime's
        # copyBuffer blit (torch's own direct-copy kernel where it copies
        # itself); it is the largest single copy dispatch of the run.  Other
        # copy-like kernels (concatenation) and per-name averages (which mix the
        # calibration copy with smaller field clones) are not used.
        best = None
        for (k, c, _), v in per_disp.items():
            if ("copyBuffer" in k or "direct_copy" in k) and (best is None or v > best[2]):
                best = (k, c, v)
        if best:
            expected_kib = calib_mib * 1024.0
            for c in ("FETCH_SIZE", "WRITE_SIZE"):
                factor[c] = expected_kib / best[2]
            print("calibration kernel: %s  measured %.1f KiB/dispatch, expected %.1f KiB -> "
                  "factor %.4f" % (_short(best[0]), best[2], expected_kib, expected_kib / best[2]))
    print("%-72s %6s %18s %18s" % ("kernel", "disp", "counter", "avg/dispatch"))
    traffic = []
    for k in sorted(acc, key=lambda k: -sum(acc[k].values())):
        nd = max(1, len(disp[k]))
        for c, v in sorted(acc[k].items()):
            f = factor.get(c, 1.0)
            print("%-72s %6d %18s %18.1f" % (_short(k), nd, c, v / nd * f))
            if c in factor:
                m = re.search(r"([A-Za-z_]\w*_kernel)\b", k)
                traffic.append("TRAFFIC\t%s\t%s\t%.0f" % (m.group(1) if m else _short(k, 40), c,
                                                           v / nd * f * 1024.0))
    for t in traffic:
        print(t)
    for k in sorted(acc, key=lambda k: -sum(acc[k].values())):
        a = acc[k]
        if a.get("SQ_ACTIVE_INST_VALU") and "SQ_THREAD_CYCLES_VALU" in a:
            m = re.search(r"([A-Za-z_]\w*_kernel)\b", k)
            print("LANES\t%s\t%.4f" % (m.group(1) if m else _short(k, 40),
                                        a["SQ_THREAD_CYCLES_VALU"] / (64.0 * a["SQ_ACTIVE_INST_VALU"])))


def main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 2
    if argv[0] == "stats":
        stats(argv[1])
    elif argv[0] == "trace":
        last = int(argv[argv.index("--last") + 1]) if "--last" in argv else 1
        trace(argv[1], last)
    elif argv[0] == "counter":
        calib = None
        if "--calib-mib" in argv:
            calib = float(argv[argv.index("--calib-mib") + 1])
        last = int(argv[argv.index("--last") + 1]) if "--last" in argv else None
        counters(argv[1], calib, last)
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
