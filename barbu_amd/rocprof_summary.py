"""Summaries of rocprofv3 SQLite output (``*_results.db``) as short text.

    python -m barbu_amd.rocprof_summary trace <results.db>   per-kernel durations
    python -m barbu_amd.rocprof_summary trace-ids <results.db>
        the same per code object kernel (kernels of one name built for
        different launch classes are told apart by id, VGPRs and LDS)
    python -m barbu_amd.rocprof_summary pmc <results.db>     per-kernel counter means
    python -m barbu_amd.rocprof_summary traffic <fetch.db> <write.db> <copy_bytes> [<json> <workload> <kernel>]
        HBM bytes per dispatch of each kernel from FETCH_SIZE / WRITE_SIZE runs,
        scaled by what the counters read for the plain copy kernel
        (ftl_copy_kernel), which reads and writes copy_bytes.  With <json>,
        also stores read + write bytes per dispatch of <kernel> under
        <workload> in that JSON file (bench.py reports it as roofline.traffic).
"""
import sqlite3
import sys
from collections import defaultdict

COPY_KERNEL = "ftl_copy_kernel"


def short(name):
    n = name.replace("(anonymous namespace)::", "")
    return n.split("(")[0]


def trace(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, duration from kernels").fetchall()
    agg = defaultdict(list)
    for name, d in rows:
        agg[short(name)].append(d / 1000.0)
    total = sum(sum(v) for v in agg.values())
    out = ["%-34s %7s %12s %10s %10s %10s %6s" % ("kernel", "calls", "total_us", "avg_us", "min_us", "max_us", "%")]
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append("%-34s %7d %12.1f %10.2f %10.2f %10.2f %6.2f" % (
            k, len(v), sum(v), sum(v) / len(v), min(v), max(v), 100.0 * sum(v) / total))
    return "\n".join(out)


def trace_ids(db):
    con = sqlite3.connect(db)
    rows = con.execute("select name, kernel_id, vgpr_count, lds_size, duration from kernels").fetchall()
    agg = defaultdict(list)
    for name, kid, vg, lds, d in rows:
        agg[(short(name), kid, vg, lds)].append(d / 1000.0)
    total = sum(sum(v) for v in agg.values())
    out = ["%-30s %4s %5s %6s %7s %12s %10s %10s %10s %6s" % ("kernel", "id", "vgpr", "lds", "calls", "total_us",
                                                               "avg_us", "min_us", "max_us", "%")]
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        out.append("%-30s %4d %5d %6d %7d %12.1f %10.2f %10.2f %10.2f %6.2f" % (
            k[0], k[1], k[2], k[3], len(v), sum(v), sum(v) / len(v), min(v), max(v), 100.0 * sum(v) / total))
    return "\n".join(out)


def pmc_means(db):
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, counter_name, value, dispatch_id from counters_collection").fetchall()
    agg = defaultdict(lambda: defaultdict(list))
    for kname, cname, val, _ in rows:
        agg[short(kname)][cname].append(val)
    return {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in agg.items()}, \
        {k: {c: len(v) for c, v in cs.items()} for k, cs in agg.items()}


def pmc(db):
    means, counts = pmc_means(db)
    out = []
    for k in sorted(means):
        for c in sorted(means[k]):
            out.append("%-34s %-18s dispatches=%4d mean=%.6g" % (k, c, counts[k][c], means[k][c]))
    return "\n".join(out)


def traffic(fetch_db, write_db, copy_bytes):
    f, _ = pmc_means(fetch_db)
    w, _ = pmc_means(write_db)
    cf = f[COPY_KERNEL]["FETCH_SIZE"] * 1024.0
    cw = w[COPY_KERNEL]["WRITE_SIZE"] * 1024.0
    kf, kw = copy_bytes / cf, copy_bytes / cw
    out = ["copy kernel: FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB for %d B read + %d B written -> scale "
           "fetch x%.4f, write x%.4f" % (cf / 1024, cw / 1024, copy_bytes, copy_bytes, kf, kw)]
    res = {}
    for k in sorted(set(f) & set(w)):
        if k == COPY_KERNEL:
            continue
        rb = f[k]["FETCH_SIZE"] * 1024.0 * kf
        wb = w[k]["WRITE_SIZE"] * 1024.0 * kw
        res[k] = (rb, wb)
        out.append("%-34s read %.4g B  write %.4g B  total %.4g B per dispatch" % (k, rb, wb, rb + wb))
    return "\n".join(out), res


def main(argv):
    if len(argv) >= 2 and argv[0] == "trace":
        print(trace(argv[1]))
    elif len(argv) >= 2 and argv[0] == "trace-ids":
        print(trace_ids(argv[1]))
    elif len(argv) >= 2 and argv[0] == "pmc":
        print(pmc(argv[1]))
    elif len(argv) >= 4 and argv[0] == "traffic":
        text, res = traffic(argv[1], argv[2], float(argv[3]))
        print(text)
        if len(argv) >= 7:
            import json
            import os
            path, workload, kernel = argv[4], argv[5], argv[6]
            data = {}
            if os.path.exists(path):
                with open(path) as f:
                    data = json.load(f)
            rb, wb = res[kernel]
            data[workload] = {"kernel": kernel, "read_bytes": rb, "write_bytes": wb, "bytes": rb + wb}
            with open(path, "w") as f:
                json.dump(data, f, indent=1)
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
