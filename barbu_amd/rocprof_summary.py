"""Summaries of rocprofv3 SQLite output (``*_results.db``) as short text.

    python -m barbu_amd.rocprof_summary trace <results.db>   per-kernel durations
    python -m barbu_amd.rocprof_summary trace-ids <results.db>
        the same per code object kernel (kernels of one name built for
        different launch classes are told apart by id, VGPRs and LDS)
    python -m barbu_amd.rocprof_summary pmc <results.db>     per-kernel counter means
    python -m barbu_amd.rocprof_summary pmc-ids <results.db> [kernel,...]
        the same per code object and workgroup count (one row per launch
        class of a batched frame), with the mean dispatch duration
    python -m barbu_amd.rocprof_summary traffic-ids <fetch.db> <write.db> <copy_bytes> <kernel,...> [<json> <workload>]
        HBM bytes per dispatch of each launch class of the named kernels, and
        their sum (a batched frame); with <json>, stores the sum
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
    if n.startswith("void "):
        n = n[5:]
    return n.split("(")[0]


def base(name):
    """Kernel name without template arguments (ftl_tiles_kernel<true> -> ftl_tiles_kernel)."""
    return name.split("<")[0]


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


def pmc_by_shape(db):
    """Counter means per dispatch, grouped by (kernel, code object id,
    workgroups): the kernels of one name built for different launch classes
    (and launched over different tile counts) are told apart.  Returns
    {key: (counter means, dispatches, mean duration us, vgprs)}."""
    con = sqlite3.connect(db)
    rows = con.execute("select kernel_name, kernel_id, grid_size, workgroup_size, vgpr_count, counter_name, value, "
                       "dispatch_id, duration from counters_collection").fetchall()
    vals = defaultdict(lambda: defaultdict(list))
    durs = defaultdict(dict)
    vg = {}
    for kn, kid, gs, wg, v, cn, val, did, d in rows:
        key = (short(kn), kid, gs // max(wg, 1))
        vals[key][cn].append(val)
        durs[key][did] = d
        vg[key] = v
    return {k: ({c: sum(x) / len(x) for c, x in cs.items()}, len(durs[k]),
                sum(durs[k].values()) / len(durs[k]) / 1000.0, vg[k]) for k, cs in vals.items()}


def pmc_ids(db, kernels=None):
    res = pmc_by_shape(db)
    names = sorted({c for v in res.values() for c in v[0]})
    out = ["%-26s %4s %8s %5s %4s %9s " % ("kernel", "id", "wgs", "vgpr", "n", "dur_us") +
           " ".join("%14s" % c for c in names)]
    for k, (means, n, dur, v) in sorted(res.items(), key=lambda kv: -kv[1][0].get("SQ_INSTS_VALU", 0)):
        if kernels and k[0] not in kernels and base(k[0]) not in kernels:
            continue
        out.append("%-26s %4d %8d %5d %4d %9.1f " % (k[0], k[1], k[2], v, n, dur) +
                   " ".join("%14.6g" % means.get(c, float("nan")) for c in names))
    return "\n".join(out)


def traffic_ids(fetch_db, write_db, copy_bytes, kernels):
    """HBM bytes per dispatch per (kernel, id, workgroups) for the named
    kernels, scaled by the copy kernel's reading; and their sum (one dispatch
    of each group: e.g. one batched frame)."""
    # The two runs are separate processes, whose code object ids may differ:
    # a launch is matched across them by kernel, workgroup count and VGPRs.
    def by_launch(res):
        return {(k[0], k[2], v[3]): (k[1], v[0]) for k, v in res.items()}
    f = by_launch(pmc_by_shape(fetch_db))
    w = by_launch(pmc_by_shape(write_db))
    cf = [v[1]["FETCH_SIZE"] for k, v in f.items() if k[0] == COPY_KERNEL][0] * 1024.0
    cw = [v[1]["WRITE_SIZE"] for k, v in w.items() if k[0] == COPY_KERNEL][0] * 1024.0
    kf, kw = copy_bytes / cf, copy_bytes / cw
    out = ["copy kernel: FETCH_SIZE %.0f KiB, WRITE_SIZE %.0f KiB for %d B read + %d B written -> scale "
           "fetch x%.4f, write x%.4f" % (cf / 1024, cw / 1024, copy_bytes, copy_bytes, kf, kw)]
    tr = tw = 0.0
    for k in sorted(set(f) & set(w)):
        if k[0] not in kernels and base(k[0]) not in kernels:
            continue
        rb = f[k][1]["FETCH_SIZE"] * 1024.0 * kf
        wb = w[k][1]["WRITE_SIZE"] * 1024.0 * kw
        tr += rb
        tw += wb
        out.append("%-26s id %3d wgs %7d  read %.4g B  write %.4g B  total %.4g B per dispatch" %
                   (k[0], f[k][0], k[1], rb, wb, rb + wb))
    out.append("sum (one dispatch of each): read %.4g B  write %.4g B  total %.4g B" % (tr, tw, tr + tw))
    return "\n".join(out), (tr, tw)


def store_traffic(path, workload, kernel, rb, wb):
    import json
    import os
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[workload] = {"kernel": kernel, "read_bytes": rb, "write_bytes": wb, "bytes": rb + wb}
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def main(argv):
    if len(argv) >= 2 and argv[0] == "pmc-ids":
        print(pmc_ids(argv[1], set(argv[2].split(",")) if len(argv) > 2 else None))
        return 0
    if len(argv) >= 5 and argv[0] == "traffic-ids":
        # traffic-ids <fetch.db> <write.db> <copy_bytes> <kernel,kernel...> [<json> <workload>]
        text, (rb, wb) = traffic_ids(argv[1], argv[2], float(argv[3]), set(argv[4].split(",")))
        print(text)
        if len(argv) >= 7:
            store_traffic(argv[5], argv[6], argv[4], rb, wb)
        return 0
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
            path, workload, kernel = argv[4], argv[5], argv[6]
            hits = [v for k, v in res.items() if k == kernel or base(k) == kernel]
            if len(hits) != 1:
                raise SystemExit("%d kernels named %s in the runs" % (len(hits), kernel))
            rb, wb = hits[0]
            store_traffic(path, workload, kernel, rb, wb)
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
