"""Summaries of rocprofv3 output for the benchmark kernels.

    python -m kites_amd.profiling <kernel-trace dir> <FETCH_SIZE dir> <WRITE_SIZE dir> \
        <bench JSON line file> <out dir> [--copy-mib 1024]

reads the CSV files rocprofv3 wrote (``--kernel-trace --stats`` and two
``--pmc`` runs of ``bench.py --copy-probe N``), and writes one short text
summary per benchmark kernel plus ``traffic.json`` (HBM bytes per launch,
which ``bench.py`` reports as ``roofline.traffic``).

    python -m kites_amd.profiling counters <out.md> <pmc dir> [<pmc dir> ...] \
        [--ktrace <kernel-trace dir>] [--labels a,b,...] [--keep 16-26,...] [--command "..."]

tabulates the SQ counters of every ``step_kernel`` dispatch of ``--pmc``
runs of the same program (one directory per pass; dispatches are matched by
order), with the share of wave time spent waiting, and the durations of a
kernel-trace run of that program when given.

FETCH_SIZE and WRITE_SIZE (KiB per dispatch) are scaled by what the same
counter reads for the probe's copy kernel, which reads and writes exactly
``--copy-mib`` MiB: factor = known KiB / counted KiB.
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


def main(argv=None):
    import sys
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "counters":
        return counters(argv[1:])
    ap = argparse.ArgumentParser()
    ap.add_argument("ktrace")
    ap.add_argument("fetch")
    ap.add_argument("write")
    ap.add_argument("bench")
    ap.add_argument("out")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--command", default="")
    a = ap.parse_args(argv)
    stats = kernel_stats(a.ktrace)
    fetch, fcopy = counter(a.fetch, "FETCH_SIZE")
    write, wcopy = counter(a.write, "WRITE_SIZE")
    known = a.copy_mib * 1024.0
    ff = known / fcopy if fcopy else None
    wf = known / wcopy if wcopy else None
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
    os.makedirs(a.out, exist_ok=True)
    traffic = {"copy_probe_mib": a.copy_mib, "fetch_factor": ff, "write_factor": wf,
               "kernels": {}}
    for k, desc in KERNELS.items():
        if k not in stats or k not in fetch:
            continue
        fkib = fetch[k][0] / fetch[k][1]
        wkib = write[k][0] / write[k][1] if k in write else 0.0
        tb = (fkib * ff + wkib * wf) * 1024.0
        roof = roofs.get(k) or {}
        alg = roof.get("bytes_per_launch")
        traffic["kernels"][k] = {"workload": work[k], "fetch_kib_per_launch": round(fkib, 1),
                                 "write_kib_per_launch": round(wkib, 1),
                                 "traffic_bytes_per_launch": int(tb),
                                 "algorithmic_bytes_per_launch": alg,
                                 "avg_ns_kernel_trace": stats[k]["avg_ns"]}
        st = stats[k]
        lines = [
            f"# {k}: {desc}",
            "",
            f"Workload: {work[k]}.",
            "",
            "Commands (from the repository root, on one MI355X):",
            "",
            "```",
            "rocprofv3 --kernel-trace --stats --output-format csv -d <kt> -o bench -- "
            "python3 -u bench.py --steps 4 --warmup 1 --no-cpu --schedules learned",
            "rocprofv3 --pmc FETCH_SIZE --output-format csv -d <fetch> -o f -- "
            "python3 -u bench.py --steps 1 --warmup 0 --no-cpu --schedules learned --copy-probe 1024",
            "rocprofv3 --pmc WRITE_SIZE --output-format csv -d <write> -o w -- "
            "python3 -u bench.py --steps 1 --warmup 0 --no-cpu --schedules learned --copy-probe 1024",
            "python3 -u bench.py --no-cpu > <bench.json>",
            "python -m kites_amd.profiling <kt> <fetch> <write> <bench.json> profiles",
            "```",
            "",
            "| quantity | value |",
            "|---|---|",
            f"| dispatches (kernel trace) | {st['calls']} |",
            f"| average duration, kernel trace | {st['avg_ns'] / 1e3:.1f} us |",
            f"| average duration per launch, bench.py HIP events | "
            f"{1e3 * roof.get('launch_ms', 0):.1f} us |",
            f"| share of GPU time in the traced run | {st['pct']:.1f} % |",
            f"| FETCH_SIZE per dispatch (raw) | {fkib:.0f} KiB |",
            f"| WRITE_SIZE per dispatch (raw) | {wkib:.0f} KiB |",
            f"| copy-probe factors (fetch, write) | {ff:.3f}, {wf:.3f} |",
            f"| HBM traffic per launch (scaled) | {tb / 1e6:.1f} MB |",
            f"| algorithmic bytes per launch | {alg / 1e6 if alg else float('nan'):.1f} MB |",
            f"| traffic / algorithmic | {tb / alg if alg else float('nan'):.2f} |",
            f"| achieved (algorithmic / time) | {roof.get('achieved')} GB/s "
            f"({100 * roof.get('frac', 0):.1f} % of 8 TB/s) |",
            f"| measured HBM rate (traffic / kernel-trace time) | "
            f"{tb / st['avg_ns']:.0f} GB/s |",
        ]
        if a.command:
            lines += ["", f"Run: `{a.command}`"]
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
