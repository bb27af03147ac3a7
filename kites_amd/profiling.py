"""Summaries of rocprofv3 output for the benchmark kernels.

    python -m kites_amd.profiling <kernel-trace dir> <read-request dir> \
        <write-request dir> <bench JSON line file> <out dir> [--copy-mib 1024]

reads the CSV files rocprofv3 wrote (``--kernel-trace --stats`` and two
``--pmc`` runs of ``bench.py --copy-probe N``: the L2's read requests by size,
TCC_EA0_RDREQ_sum / TCC_EA0_RDREQ_32B_sum / TCC_BUBBLE_sum, then its write
requests, TCC_EA0_WRREQ_sum / TCC_EA0_WRREQ_64B_sum), and writes one short
text summary per benchmark kernel plus ``traffic.json`` (HBM bytes per
launch, which ``bench.py`` reports as ``roofline.traffic``).

    python -m kites_amd.profiling counters <out.md> <pmc dir> [<pmc dir> ...] \
        [--ktrace <kernel-trace dir>] [--labels a,b,...] [--keep 16-26,...] [--command "..."]

tabulates the SQ counters of every ``step_kernel`` dispatch of ``--pmc``
runs of the same program (one directory per pass; dispatches are matched by
order), with the share of wave time spent waiting, and the durations of a
kernel-trace run of that program when given.

Bytes follow the request sizes (32, 64 or 128 B per read request, 32 or 64
B per write request); the probe's copy kernel, which reads and writes
exactly ``--copy-mib`` MiB, shows whether the 128-byte read counter sees
its requests.  If it does not, every read request that is not 32 B is
taken as 128 B (an upper bound) and the summary says so.
"""
import argparse
import csv
import glob
import json
import os

KERNELS = {
    "step_kernel": "single-source iteration kernel (one launch per iteration)",
    "batch_step_kernel": "64-source K-wide iteration kernel",
    "bitbfs_step_kernel": "256-source bit-parallel BFS kernel",
}


def short_name(full):
    for k in ("batch_step_kernel", "bitbfs_step_kernel", "step_kernel"):
        if k + "<" in full or k + "(" in full:
            return k
    if "copyBuffer" in full:
        return "copy"
    return None


def one(path_glob):
    files = glob.glob(path_glob)
    if not files:
        raise FileNotFoundError(path_glob)
    return files[0]


def kernel_stats(d):
    out = {}
    for r in csv.DictReader(open(one(os.path.join(d, "*kernel_stats.csv")))):
        k = short_name(r["Name"])
        if k in KERNELS:
            out[k] = dict(calls=int(r["Calls"]), avg_ns=float(r["AverageNs"]),
                          total_ns=float(r["TotalDurationNs"]), pct=float(r["Percentage"]))
    return out


def counter(d, name):
    per = {}
    copy_max = 0.0
    for r in csv.DictReader(open(one(os.path.join(d, "*counter_collection.csv")))):
        if r["Counter_Name"] != name:
            continue
        k = short_name(r["Kernel_Name"])
        v = float(r["Counter_Value"])
        if k == "copy":
            copy_max = max(copy_max, v)
        elif k in KERNELS:
            s = per.setdefault(k, [0.0, 0])
            s[0] += v
            s[1] += 1
    return per, copy_max


def _step_dispatches(directory):
    """{counter: value} per step_kernel dispatch, in dispatch order."""
    per = {}
    for path in glob.glob(os.path.join(directory, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if short_name(r["Kernel_Name"]) != "step_kernel":
                    continue
                d = per.setdefault(int(r["Dispatch_Id"]), {})
                d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return [per[k] for k in sorted(per)]


def _step_durations(directory):
    rows = []
    for path in glob.glob(os.path.join(directory, "**", "*kernel_trace.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if short_name(r["Kernel_Name"]) == "step_kernel":
                    rows.append((int(r["Start_Timestamp"]),
                                 (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3))
    return [d for _, d in sorted(rows)]


def counters(argv):
    ap = argparse.ArgumentParser(prog="kites_amd.profiling counters")
    ap.add_argument("out")
    ap.add_argument("pmc", nargs="+")
    ap.add_argument("--ktrace", default="")
    ap.add_argument("--labels", default="")
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--keep", default="",
                    help="dispatch ranges to list, e.g. 16-26,48-58 (default: all)")
    a = ap.parse_args(argv)
    keep = set()
    for part in filter(None, a.keep.split(",")):
        lo, _, hi = part.partition("-")
        keep.update(range(int(lo), int(hi or lo) + 1))
    passes = [_step_dispatches(d) for d in a.pmc]
    count = min(len(p) for p in passes)
    dur = _step_durations(a.ktrace) if a.ktrace else []
    labels = a.labels.split(",") if a.labels else []
    names = []
    for p in passes:
        for k in (p[0] if p else {}):
            if k not in names:
                names.append(k)
    head = ["dispatch", "run"] + (["us"] if dur else []) + [n.replace("SQ_", "") for n in names]
    if "SQ_WAIT_ANY" in names and "SQ_WAVE_CYCLES" in names:
        head.append("wait share")
    lines = ["| " + " | ".join(head) + " |", "|" + "---|" * len(head)]
    per_run = (count // len(labels)) if labels else count
    for i in range(count):
        if keep and i not in keep:
            continue
        merged = {}
        for p in passes:
            merged.update(p[i])
        run = labels[i // per_run] if labels and i // per_run < len(labels) else ""
        row = [str(i), run] + ([f"{dur[i]:.1f}" if i < len(dur) else ""] if dur else [])
        row += [f"{merged.get(n, 0):.3g}" for n in names]
        if "wait share" in head:
            wc = merged.get("SQ_WAVE_CYCLES", 0)
            row.append(f"{merged.get('SQ_WAIT_ANY', 0) / wc:.2f}" if wc else "")
        lines.append("| " + " | ".join(row) + " |")
    text = ["# step_kernel: SQ counters per iteration", ""]
    if a.command:
        text += ["Commands (from the repository root, on one MI355X):", "", "```"] + a.command + ["```", ""]
    text += lines
    with open(a.out, "w") as f:
        f.write("\n".join(text) + "\n")
    print("\n".join(lines))


RD = ("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_BUBBLE_sum")
WR = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")


def request_bytes(c):
    """HBM-side bytes from the L2's request counters of one dispatch (gfx950:
    TCC_EA0_RDREQ counts read requests of every size, RDREQ_32B the 32-byte
    ones, TCC_BUBBLE the 128-byte ones, the rest are 64 bytes; WRREQ counts
    write requests, WRREQ_64B the 64-byte ones, the rest are 32 bytes).
    Returns (read, write, read if every request that is not 32 B were 128 B)."""
    r, r32, b = (c.get(k, 0.0) for k in RD)
    w, w64 = (c.get(k, 0.0) for k in WR)
    read = 32 * r32 + 128 * b + 64 * max(0.0, r - r32 - b)
    write = 32 * max(0.0, w - w64) + 64 * w64
    read_hi = 32 * r32 + 128 * max(0.0, r - r32)
    return read, write, read_hi


def per_kernel_counters(dirs):
    """{kernel: {counter: average per dispatch}} over the --pmc runs in
    `dirs`, plus the same for the copy probe's kernel (largest dispatch)."""
    out, copy = {}, {}
    for d in dirs:
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                rows = list(csv.DictReader(f))
            per = {}
            for r in rows:
                k = short_name(r["Kernel_Name"])
                if k is None:
                    continue
                key = (k, int(r["Dispatch_Id"]))
                per.setdefault(key, {})[r["Counter_Name"]] = float(r["Counter_Value"])
            sums = {}
            for (k, _), c in per.items():
                if k == "copy":
                    for name, v in c.items():
                        copy[name] = max(copy.get(name, 0.0), v)
                    continue
                s = sums.setdefault(k, [{}, 0])
                s[1] += 1
                for name, v in c.items():
                    s[0][name] = s[0].get(name, 0.0) + v
            for k, (tot, cnt) in sums.items():
                dst = out.setdefault(k, {})
                for name, v in tot.items():
                    dst[name] = v / cnt
    return out, copy


def main(argv=None):
    import sys
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "counters":
        return counters(argv[1:])
    ap = argparse.ArgumentParser()
    ap.add_argument("ktrace")
    ap.add_argument("rd", help="--pmc run of the read-request counters")
    ap.add_argument("wr", help="--pmc run of the write-request counters")
    ap.add_argument("bench")
    ap.add_argument("out")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--command", action="append", default=[])
    a = ap.parse_args(argv)
    stats = kernel_stats(a.ktrace)
    cnt, copy = per_kernel_counters([a.rd, a.wr])
    known = a.copy_mib * float(1 << 20)
    crd, cwr, crd_hi = request_bytes(copy) if copy else (0.0, 0.0, 0.0)
    bench = None
    for line in open(a.bench):
        if line.startswith("{"):
            bench = json.loads(line)
    roofs = {"step_kernel": bench.get("roofline"),
             "batch_step_kernel": (bench.get("batch64") or {}).get("roofline"),
             "bitbfs_step_kernel": (bench.get("bfs256") or {}).get("roofline")}
    work = {"step_kernel": bench["config"]["workload"],
            "batch_step_kernel": (bench.get("batch64") or {}).get("workload"),
            "bitbfs_step_kernel": (bench.get("bfs256") or {}).get("workload")}
    # The copy probe moves `known` bytes each way: when its resolved read
    # bytes fall short (128-byte requests the 128-byte counter missed), the
    # upper bound, every request that is not 32 B taken as 128 B, is used.
    use_hi = bool(copy) and crd < 0.9 * known <= crd_hi * 1.05
    os.makedirs(a.out, exist_ok=True)
    traffic = {"copy_probe_mib": a.copy_mib, "copy_probe": {
        "counters": copy, "read_bytes": crd, "read_bytes_if_128B": crd_hi, "write_bytes": cwr,
        "known_bytes": known}, "non_32B_reads_taken_as_128B": use_hi, "kernels": {}}
    cmds = a.command or ["(commands not recorded)"]
    for k, desc in KERNELS.items():
        if k not in stats or k not in cnt:
            continue
        c = cnt[k]
        rd, wr, rd_hi = request_bytes(c)
        rb = rd_hi if use_hi else rd
        tb = rb + wr
        roof = roofs.get(k) or {}
        alg = roof.get("bytes_per_launch")
        st = stats[k]
        traffic["kernels"][k] = {"workload": work[k], "counters_per_dispatch": c,
                                 "read_bytes_per_launch": int(rb),
                                 "read_bytes_64B_reading": int(rd),
                                 "read_bytes_128B_reading": int(rd_hi),
                                 "write_bytes_per_launch": int(wr),
                                 "traffic_bytes_per_launch": int(tb),
                                 "algorithmic_bytes_per_launch": alg,
                                 "avg_ns_kernel_trace": st["avg_ns"]}
        r, r32, b = (c.get(x, 0.0) for x in RD)
        w, w64 = (c.get(x, 0.0) for x in WR)
        lines = [
            f"# {k}: {desc}",
            "",
            f"Workload: {work[k]}.",
            "",
            "Commands (from the repository root, on one MI355X):",
            "",
            "```"] + cmds + [
            "```",
            "",
            "| quantity | value |",
            "|---|---|",
            f"| dispatches (kernel trace) | {st['calls']} |",
            f"| average duration, kernel trace | {st['avg_ns'] / 1e3:.1f} us |",
            f"| average duration per launch, bench.py HIP events | "
            f"{1e3 * roof.get('launch_ms', 0):.1f} us |",
            f"| share of GPU time in the traced run | {st['pct']:.1f} % |",
            f"| read requests per dispatch: all / 32 B / 128 B (TCC_BUBBLE) | "
            f"{r:.4g} / {r32:.4g} / {b:.4g} |",
            f"| write requests per dispatch: all / 64 B | {w:.4g} / {w64:.4g} |",
            f"| read bytes, 64-B reading of the rest / all 128 B | {rd / 1e6:.1f} / "
            f"{rd_hi / 1e6:.1f} MB |",
            f"| copy probe ({a.copy_mib} MiB each way): read bytes resolved / if 128 B, write | "
            f"{crd / 2**20:.0f} / {crd_hi / 2**20:.0f} / {cwr / 2**20:.0f} MiB |",
            f"| HBM traffic per launch (read{' (128-B reading)' if use_hi else ''} + write) | "
            f"{tb / 1e6:.1f} MB |",
            f"| algorithmic bytes per launch | {alg / 1e6 if alg else float('nan'):.1f} MB |",
            f"| traffic / algorithmic | {tb / alg if alg else float('nan'):.2f} |",
            f"| achieved (algorithmic / time) | {roof.get('achieved')} GB/s "
            f"({100 * roof.get('frac', 0):.1f} % of 8 TB/s) |",
            f"| measured HBM rate (traffic / kernel-trace time) | "
            f"{tb / st['avg_ns']:.0f} GB/s |",
        ]
        name = {"step_kernel": "step_kernel_rmat22_sssp1.md",
                "batch_step_kernel": "batch_step_kernel_rmat22_sssp64.md",
                "bitbfs_step_kernel": "bitbfs_step_kernel_rmat22_bfs256.md"}[k]
        with open(os.path.join(a.out, name), "w") as f:
            f.write("\n".join(lines) + "\n")
    with open(os.path.join(a.out, "traffic.json"), "w") as f:
        json.dump(traffic, f, indent=1)
        f.write("\n")
    print(json.dumps(traffic, indent=1))


if __name__ == "__main__":
    main()
