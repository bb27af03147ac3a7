"""Benchmark: SSSP GTEPS on R-MAT scale 22 (edge factor 16) on one MI355X.

Workloads (all on the same seeded graph: R-MAT a=.57 b=.19 c=.19, 2^22
vertices, 2^26 edges drawn, self-loops and duplicates removed, weights
U[1,255]; generated and uploaded before any timed region):

  single   one-source SSSP from the max-out-degree vertex.  A step is one
           run of the engine (device-side init of values and frontier, then
           the fixpoint).  Timed for the plain schedule and for the learned
           schedule (kites_amd/models/scheduler.npz); `value` is the learned
           schedule's GTEPS, the plain one is reported beside it.  The
           schedules' timed steps are interleaved (one step of each in
           turn), so a drift of the device's speed reaches them alike.
  batch64  64-source SSSP in one K-wide run (64 sources drawn with a fixed
           seed among vertices with out-edges).  A step is one batch run.
  bfs256   256-source bit-parallel BFS in one run.
  ingest   the graph build itself: the benchmark graph as a shuffled edge
           list, built on the GPU and on the host (all cores), both checked
           equal to the generator's arrays.
  highdiam the 2048 x 2048 grid (a high-diameter graph: ~4,000 iterations
           per run): SSSP and BFS with and without the one-workgroup loop
           for small frontiers, and near-far SSSP; time per iteration.

GTEPS = Graph500 numerator (out-edges of reached vertices, summed over
sources) / wall time per step.  The roofline objects use algorithmic bytes
(the per-unit formulas of alg_bytes_single / alg_bytes_batch /
alg_bytes_bfs256 below, counted from the engine's own per-iteration
statistics) over the kernel's device time measured with HIP events on the
engine's stream; peak 8 TB/s HBM.  `cpu_baseline` times the oracle's
multi-core frontier Bellman-Ford on this host's cores, and its single-core
Gauss-Seidel restatement of the reference's CPU path, on the same graph and
source, and asserts identical distances.

Prints one JSON line (driver contract) plus progress on stderr.
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
METRIC = "SSSP GTEPS at RMAT-22 (1-src & 64-src), learned vs plain schedule; HBM %"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=int, default=16)
    ap.add_argument("--seed", type=int, default=22)
    ap.add_argument("--workload", default="all",
                    help="all (= single,batch64,bfs256,ingest,highdiam) | single | batch64 | bfs256 | "
                         "ingest | graphs1000 | variants | nearfar | highdiam (comma list; variants: one "
                         "SSSP run per forced edge-balanced variant, direct launches, for "
                         "counter profiles; nearfar: SSSP time and edges walked per near-far "
                         "delta; highdiam: 2048^2 grid SSSP/BFS with and without the "
                         "one-workgroup loop)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch64-shape", default="",
                    help="edges_per_item,block,unroll for batch64 instead of the autotuner's "
                         "(profiling runs: no tuning launches in the trace)")
    ap.add_argument("--bfs256-shape", default="",
                    help="edges_per_item,block for bfs256 instead of the autotuner's")
    ap.add_argument("--schedules", default="plain,plain_tuned,push_only,learned",
                    help="single-source schedules to time: plain (threshold 14), plain_tuned "
                         "(threshold tuned on the training graphs), push_only (edge-balanced "
                         "push every iteration), learned (the reported value; 'learned' alone "
                         "makes a kernel trace hold only the launches the roofline describes)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="MiB copied by one torch copy kernel first (PMC calibration)")
    return ap.parse_args()


def alg_bytes_single(stats, n):
    """Algorithmic bytes of the single-source iterations from the engine's
    counters.  Push (the path's model; the one-workgroup loop's iterations
    are pushes): 16 per queue entry (id, R_out pair, V[u]) + 12 per out-edge
    (C_out, E_out, 4-B read-modify-write of V[dst]) + 4 per next-queue
    entry.  CLASSIC: 12 per active vertex and per in-edge plus the bitmap
    and 8 per update.  Pull: 12 n + 4 per in-edge + 8 per gathered source +
    12 per update.  INIT_LB: 4 (n + 1) + 12 per gathered
    in-edge + 12 per update."""
    total = 0
    words = 2 * ((n + 7) // 8)          # bitmap read + clear, bytes
    for s in stats:
        v, upd = s["variant"], s["count"]
        if v == 0:      # classic
            total += words + 12 * s["active"] + 12 * s["work_edges"] + 8 * upd
        elif v <= 4 or v == 11:    # push (11: a push of the one-workgroup loop)
            total += 16 * s["active"] + 12 * s["work_edges"] + 4 * upd
        elif v <= 8:    # pull: R + V per vertex; C per in-edge; E + V[src] per gather
            total += 12 * n + 4 * s["work_edges"] + 8 * s["gathers"] + 12 * upd + words
        elif v == 9:    # init
            total += 4 * (n + 1) + 12 * s["gathers"] + 12 * upd + words
        else:           # near-far split: far entry id + V read, kept entry written
            total += 8 * s["active"] + 4 * s["active"] + 4 * upd
    return total


def alg_bytes_batch(stats, n, m):
    """K-wide batch: per iteration every in-edge's source id and weight
    (4 + 4 B) and its source's frontier bit (1/8 B), the whole 256-B row of
    every active in-edge's source, the 256-B row of every flushed
    destination and 4 B per moved value, and the frontier-ring clear
    (1/8 B per vertex)."""
    total = 0
    for s in stats:
        total += 8 * m + m // 8 + 256 * s["active_edges"] + 256 * s["flushes"] \
            + 4 * s["moved"] + n // 8
    return total


def alg_bytes_bfs256(stats, n, m, sources=256):
    """Bit-parallel BFS.  A pulled level: every in-edge's source id (4 B)
    and summary bit (1/8 B), each active edge's 32-B mask.  A top-down
    level: the frontier summary (1/8 B per vertex), each frontier vertex's
    out-offsets and 32-B mask (40 B), each of its out-edges' destination id
    and done bit (4 + 1/8 B), and each active edge's 32-B visited row.
    Either way: 32-B visited read and new-mask write per updated row, the
    clear of the masks written two iterations before (32 B per row updated
    then; summed over the run, 32 B per updated row) with the summary scan
    (1/8 B per vertex); byte levels are written once per reached pair
    (1 B).  The final u32 expansion is a kernel of its own."""
    total = 0
    for i, s in enumerate(stats):
        if s.get("push"):
            frontier = stats[i - 1]["updated"] if i else sources
            walked = stats[i - 1]["out_edges"] if i else s["active_edges"]
            total += n // 8 + 40 * frontier + 4 * walked + walked // 8 + 32 * s["active_edges"]
        else:
            total += 4 * m + m // 8 + 32 * s["active_edges"]
        total += 64 * s["updated"] + 32 * s["updated"] + n // 8 + s["new_pairs"]
    return total


def profiled_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    summary (profiles/traffic.json: FETCH_SIZE + WRITE_SIZE per dispatch,
    scaled by the copy-kernel probe), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)
        return t["kernels"][kernel]["traffic_bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def roofline(bytes_alg, device_ms, launches, kernel):
    achieved = bytes_alg / (device_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(kernel), "kernel": kernel,
            "bytes_per_launch": int(bytes_alg / max(1, launches)),
            "launch_ms": round(device_ms / max(1, launches), 5), "launches": launches}


def time_interleaved(fns, steps, warmup):
    """Times several step functions with their timed steps interleaved
    (step k of each in turn, after every function's warmup), so that a
    drift of the device's speed during the run reaches them alike; each
    step is synchronous.  Returns {name: (wall ms per step, median device
    ms)}."""
    for fn in fns.values():
        for _ in range(warmup):
            fn()
    walls = {k: 0.0 for k in fns}
    devs = {k: [] for k in fns}
    for _ in range(steps):
        for k, fn in fns.items():
            t0 = time.perf_counter()
            devs[k].append(fn())
            walls[k] += time.perf_counter() - t0
    return {k: (walls[k] * 1e3 / max(1, steps), float(np.median(devs[k])) if devs[k] else 0.0)
            for k in fns}


def time_steps(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.append(fn())
    wall_ms = (time.perf_counter() - t0) * 1e3 / max(1, steps)
    return wall_ms, float(np.median(dev)) if dev else 0.0


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; sharding a graph or a batch "
                          "over several GPUs is not implemented", "n_gpus": a.gpus}))
        sys.exit(2)
    import kites_amd as K
    from kites_amd import scheduler as S

    if a.copy_probe:
        import torch
        nbytes = a.copy_probe << 20
        x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        log(f"copy probe: {nbytes} bytes read and written by one copy kernel")

    work = ({"single", "batch64", "bfs256", "ingest", "highdiam"} if a.workload == "all"
            else set(a.workload.split(",")))
    t0 = time.time()
    h = K.HostGraph.rmat(a.scale, a.edge_factor, a.seed, wmin=1, wmax=255)
    R, C, E, oR, oC, oE = h.arrays()
    n, m = len(R) - 1, len(C)
    outdeg = np.diff(oR.astype(np.int64))
    src = int(np.argmax(outdeg))
    log(f"graph n={n} m={m} src={src} outdeg={outdeg[src]} gen {time.time() - t0:.1f}s")
    d = K.DeviceGraph(h)
    out = {"metric": METRIC, "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
           "data": "synthetic (seeded R-MAT generated on the host; no dataset)"}
    cfg = {"workload": f"SSSP 1-source, R-MAT scale {a.scale} edge factor {a.edge_factor} "
                       f"seed {a.seed}, weights U[1,255], source = max out-degree vertex",
           "n": n, "m": m, "source": src}

    if "single" in work:
        w = S.load_model()
        d.set_model(w)
        res = {}
        dist = None
        kwmap = dict(S.baseline_schedules())
        kwmap["learned"] = dict(sched="learned")
        scheds = [x for x in a.schedules.split(",") if x in kwmap]
        if "learned" not in scheds:
            raise SystemExit("--schedules must include learned (the reported value)")
        def make_step(kw):
            def step():
                d.init("sssp", src)
                return d.run("sssp", **kw)["ms"]
            return step
        # the schedules' timed steps interleaved (one step of each in turn)
        timed = time_interleaved({s: make_step(kwmap[s]) for s in scheds}, a.steps, a.warmup)
        for sched in scheds:
            kw = kwmap[sched]
            wall, dev = timed[sched]
            d.init("sssp", src)
            d.run("sssp", **kw)
            vals = d.values()
            if dist is None:
                dist = vals
            assert np.array_equal(vals, dist), f"{sched} result differs"
            d.init("sssp", src)
            info = d.run("sssp", keep_stats=True, **kw)
            st = d.stats()
            mix = {}
            for s in st:
                mix[K.VARIANTS[s["variant"]]] = mix.get(K.VARIANTS[s["variant"]], 0) + 1
            res[sched] = dict(wall_ms=wall, dev_ms=dev, iters=info["iterations"], mix=mix,
                              bytes=alg_bytes_single(st, n), launches=info["launches"]
                              if "launches" in info else None,
                              work=int(sum(s["work_edges"] for s in st)))
        reached = dist != K.INF
        teps_edges = int(outdeg[reached].sum())
        for sched in res:
            res[sched]["gteps"] = teps_edges / (res[sched]["wall_ms"] * 1e-3) / 1e9
            # work-TEPS: edges the iterations actually walked (pull: in-edges
            # scanned; push: out-edges expanded), per second of device time
            res[sched]["work_gteps"] = res[sched]["work"] / (res[sched]["dev_ms"] * 1e-3) / 1e9
            log(f"{sched}: {res[sched]['wall_ms']:.3f} ms/step wall, {res[sched]['dev_ms']:.3f} ms "
                f"device, {res[sched]['iters']} iterations, {res[sched]['gteps']:.2f} GTEPS, "
                f"mix {res[sched]['mix']}")
        L = res["learned"]
        # launches: the engine replays chunks of 16 launches of one kernel
        launches = 16 * ((L["iters"] + 15) // 16)
        out.update({"value": round(L["gteps"], 4), "unit": "GTEPS",
                    "ms_per_step": round(L["wall_ms"], 4)})
        cfg.update({"schedule": "learned",
                    "reached_edges": teps_edges, "iterations": L["iters"],
                    "variant_mix": L["mix"], "device_ms": round(L["dev_ms"], 4),
                    "work_edges": L["work"], "work_gteps": round(L["work_gteps"], 4)})
        for name in ("plain", "plain_tuned", "push_only"):
            P = res.get(name)
            cfg[name] = None if P is None else {
                "gteps": round(P["gteps"], 4), "ms_per_step": round(P["wall_ms"], 4),
                "device_ms": round(P["dev_ms"], 4), "iterations": P["iters"],
                "variant_mix": P["mix"], "work_edges": P["work"],
                "work_gteps": round(P["work_gteps"], 4), "options": kwmap[name]}
            cfg["learned_over_" + name] = None if P is None else round(L["gteps"] / P["gteps"], 4)
        out["roofline"] = roofline(L["bytes"], L["dev_ms"], launches, "step_kernel")

        cpu = None
        if not a.no_cpu:
            from oracle import liboracle as O
            V0, F0 = O.sssp_init(n, src, R, C)
            # median of 3 runs each: the multi-core frontier port on every host
            # core, and the single-core Gauss-Seidel restatement of the
            # reference's CPU path
            tms, tgs = [], []
            for _ in range(3):
                # the graph's own out-CSR is passed in: only the relaxation
                # rounds are timed, not a graph build
                tc = time.perf_counter()
                Vc, rounds, used = O.frontier_omp_csr(R, C, E, oR, oC, oE, V0, F0, O.SSSP_SAT,
                                                      threads=a.cpu_threads)
                tms.append(time.perf_counter() - tc)
                assert np.array_equal(Vc, dist), "GPU result differs from the CPU baseline"
            for _ in range(3):
                tc = time.perf_counter()
                Vg, sweeps = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
                tgs.append(time.perf_counter() - tc)
                assert np.array_equal(Vg, dist), "GPU result differs from the reference CPU path"
            tcpu, tgs1 = float(np.median(tms)), float(np.median(tgs))
            cpu = {"value": round(teps_edges / tcpu / 1e9, 5), "unit": "GTEPS", "cores": used,
                   "kind": "port",
                   "sample": f"one full single-source SSSP of the benchmark graph from the same "
                             f"source, median of 3 ({tcpu:.2f} s, oracle multi-core frontier "
                             f"Bellman-Ford, {rounds} rounds), distances identical to the GPU's",
                   "single_core": {"value": round(teps_edges / tgs1 / 1e9, 5), "unit": "GTEPS",
                                   "cores": 1, "seconds": round(tgs1, 3), "sweeps": sweeps,
                                   "sample": "the oracle's Gauss-Seidel restatement of the "
                                             "reference CPU path (sanity_check.cuh:21-47), "
                                             "same graph and source, median of 3"}}
            log(f"cpu {cpu['value']:.4f} GTEPS on {used} cores, "
                f"{cpu['single_core']['value']:.4f} GTEPS Gauss-Seidel on 1 core")
        out["cpu_baseline"] = cpu

    if "batch64" in work:
        cand = np.flatnonzero(outdeg > 0)
        src64 = np.random.default_rng(64).choice(cand, 64, replace=False).astype(np.uint32)
        # work-item shape from the autotuner (every shape checked identical),
        # or the one given
        if a.batch64_shape:
            shape64 = dict(zip(("edges_per_item", "block", "unroll"),
                               (int(x) for x in a.batch64_shape.split(","))))
        else:
            tb, _, _ = d.batch_autotune("sssp", src64, repeats=2)
            shape64 = {k: tb[k] for k in ("edges_per_item", "block", "unroll")}
        log(f"batch64 shape {shape64}")

        def bstep():
            return d.batch_run("sssp", src64, **shape64)["ms"]
        wall, dev = time_steps(bstep, max(1, a.steps // 2), max(1, a.warmup // 2))
        rows, info = d.batch("sssp", src64, keep_stats=True, **shape64)
        st = d.batch_stats()
        reached = rows != K.INF
        teps = int(sum(outdeg[reached[:, k]].sum() for k in range(64)))
        byts = alg_bytes_batch(st, n, m)
        launches = 8 * ((info["iterations"] + 7) // 8)
        b = {"workload": f"SSSP from 64 sources in one K-wide run, R-MAT scale {a.scale} "
                         f"edge factor {a.edge_factor} seed {a.seed}, weights U[1,255]",
             "gteps": round(teps / (wall * 1e-3) / 1e9, 3), "ms_per_step": round(wall, 3),
             "device_ms": round(dev, 3), "iterations": info["iterations"],
             "reached_edges": teps, "sources": 64, "source_seed": 64, "shape": shape64,
             "roofline": roofline(byts, dev, launches, "batch_step_kernel")}
        b["hbm_frac"] = b["roofline"]["frac"]
        # the stated alternative design: per-source frontier masks besides the
        # one-bit frontier (same rows, timed the same way)
        def mstep():
            return d.batch_run("sssp", src64, per_source_masks=True)["ms"]
        mwall, mdev = time_steps(mstep, max(1, a.steps // 2), max(1, a.warmup // 2))
        mrows, _ = d.batch("sssp", src64, per_source_masks=True)
        assert np.array_equal(mrows, rows), "per-source masks change the batch result"
        b["per_source_masks"] = {"ms_per_step": round(mwall, 3), "device_ms": round(mdev, 3),
                                 "gteps": round(teps / (mwall * 1e-3) / 1e9, 3)}
        log(f"batch64 with per-source masks: {mwall:.2f} ms/step (one-bit frontier {wall:.2f})")
        out["batch64"] = b
        log(f"batch64: {wall:.2f} ms/step, {b['gteps']} GTEPS, {b['roofline']['achieved']} GB/s "
            f"algorithmic ({100 * b['hbm_frac']:.1f} % of HBM peak)")
        if "single" not in work:
            out.update({"value": b["gteps"], "unit": "GTEPS", "ms_per_step": b["ms_per_step"]})
            cfg["workload"] = b["workload"]

    if "bfs256" in work:
        cand = np.flatnonzero(outdeg > 0)
        src256 = np.random.default_rng(256).choice(cand, 256, replace=False).astype(np.uint32)
        # work-item shape from the autotuner (every shape's levels checked
        # identical through a device digest)
        if a.bfs256_shape:
            shape256 = dict(zip(("edges_per_item", "block"),
                                (int(x) for x in a.bfs256_shape.split(","))))
        else:
            tb, _, _ = d.bfs256_autotune(src256, repeats=2)
            shape256 = {k: tb[k] for k in ("edges_per_item", "block")}
        log(f"bfs256 shape {shape256}")

        def fstep():
            return d.bfs256(src256, copy=False, **shape256)[1]["ms"]
        wall, dev = time_steps(fstep, max(1, a.steps // 2), max(1, a.warmup // 2))
        _, info = d.bfs256(src256, copy=False, keep_stats=True, **shape256)
        st = d.bfs256_stats()
        pairs = sum(s["new_pairs"] for s in st)
        byts = alg_bytes_bfs256(st, n, m)
        launches = 8 * ((info["iterations"] + 7) // 8)
        f = {"workload": f"BFS from 256 sources, bit-parallel, R-MAT scale {a.scale} "
                         f"edge factor {a.edge_factor} seed {a.seed}",
             "ms_per_step": round(wall, 3), "device_ms": round(dev, 3),
             "iterations": info["iterations"], "sources": 256, "reached_pairs": pairs,
             "shape": shape256,
             "roofline": roofline(byts, dev, launches, "bitbfs_step_kernel")}
        out["bfs256"] = f
        log(f"bfs256: {wall:.2f} ms/step, {f['roofline']['achieved']} GB/s algorithmic")
        if "value" not in out:
            out.update({"value": round(wall, 4), "unit": "ms", "higher_is_better": False,
                        "ms_per_step": round(wall, 4)})
    if "ingest" in work:
        # GPU graph build: the benchmark graph as an edge list in a shuffled
        # order, built on the GPU (two stable radix sorts) and on the host
        # (all cores); both must give the generator's arrays exactly.
        rows = np.repeat(np.arange(n, dtype=np.uint32), np.diff(R.astype(np.int64)))
        perm = np.random.default_rng(5).permutation(m)
        es, ed, ew = C[perm], rows[perm], E[perm]
        del rows, perm
        gms, gwall = [], []
        for _ in range(3):
            tg = time.perf_counter()
            gi, bms = K.DeviceGraph.from_edges(n, es, ed, ew)
            gwall.append((time.perf_counter() - tg) * 1e3)
            gms.append(bms)
            if len(gms) < 3:
                gi.close()
        got = gi.arrays()
        assert all(np.array_equal(x, y) for x, y in zip(got[:6], (R, C, E, oR, oC, oE))), \
            "GPU-built graph differs from the generator's"
        gi.close()
        hms = []
        for _ in range(3 if not a.no_cpu else 1):
            th = time.perf_counter()
            hi = K.HostGraph.from_edges(n, es, ed, ew)
            hms.append((time.perf_counter() - th) * 1e3)
        assert all(np.array_equal(x, y) for x, y in zip(hi.arrays(), (R, C, E, oR, oC, oE)))
        del hi
        cores = len(os.sched_getaffinity(0))
        ig = {"workload": f"graph build from an edge list of {m} edges over {n} vertices "
                          f"(R-MAT scale {a.scale} edge factor {a.edge_factor} seed {a.seed}, "
                          f"shuffled): in- and out-CSR, non-empty list, out-edge bitmap",
              "gpu_build_ms": round(float(np.median(gms)), 3),
              "gpu_wall_ms_incl_upload": round(float(np.median(gwall)), 3),
              "host_build_ms": round(float(np.median(hms)), 3), "host_cores": cores,
              "medges_per_s_gpu": round(m / (float(np.median(gms)) * 1e-3) / 1e6, 1),
              "identical_to_generator": True}
        out["ingest"] = ig
        log(f"ingest: GPU {ig['gpu_build_ms']} ms ({ig['gpu_wall_ms_incl_upload']} ms with the "
            f"upload), host {ig['host_build_ms']} ms on {cores} cores")

    if "variants" in work:
        # One SSSP run per forced edge-balanced variant (every iteration the
        # same variant), direct launches, after one warm-up run each: the
        # workload of the per-iteration counter profiles.
        vres, ref = {}, None
        for name, kw in (("push_lb", dict(forced_variant="push_lb")),
                         ("pull_lb", dict(forced_variant="pull_lb")),
                         ("pull_lb_split", dict(forced_variant="pull_lb", split_pull=True))):
            for _ in range(2):
                d.init("sssp", src)
                info = d.run("sssp", sched="forced", use_graph=False, **kw)
            vals = d.values()
            ref = vals if ref is None else ref
            assert np.array_equal(vals, ref), f"{name} differs"
            vres[name] = {"device_ms": round(info["ms"], 4), "iterations": info["iterations"]}
            log(f"variants: {name} {info['ms']:.3f} ms, {info['iterations']} iterations")
        out["variants"] = vres

    if "nearfar" in work:
        # Near-far push on the benchmark graph: device time (best of 5), the
        # edges the iterations walked and the split iterations, per delta,
        # for the plain, forced edge-balanced push and learned schedules.
        d.set_model(S.load_model())
        ref = None
        rows = []
        for sched in ("push_lb", "plain", "learned"):
            kw = (dict(sched="forced", forced_variant="push_lb") if sched == "push_lb"
                  else dict(sched=sched))
            for delta in (0, 16, 32, 64, 128, 256, 512, 1024):
                best = None
                for _ in range(5):
                    d.init("sssp", src)
                    ms = d.run("sssp", delta=delta, **kw)["ms"]
                    best = ms if best is None else min(best, ms)
                vals = d.values()
                if ref is None:
                    ref = vals
                assert np.array_equal(vals, ref), f"near-far {sched} delta {delta} differs"
                d.init("sssp", src)
                info = d.run("sssp", delta=delta, keep_stats=True, **kw)
                st = d.stats()
                walked = int(sum(x["work_edges"] for x in st))
                splits = sum(1 for x in st if K.VARIANTS[x["variant"]] == "split")
                rows.append({"schedule": sched, "delta": delta, "device_ms": round(best, 4),
                             "iterations": info["iterations"], "splits": splits,
                             "edges_walked": walked})
                log(f"nearfar {sched} delta {delta}: {best:.3f} ms, {info['iterations']} "
                    f"iterations ({splits} splits), {walked} edges walked")
        out["nearfar"] = rows

    if "highdiam" in work:
        # High-diameter graph: the 2048 x 2048 4-neighbour grid (weights
        # U[1,100], seed 201), SSSP and BFS from corner 0, plain and learned
        # schedules, without the one-workgroup loop for small push frontiers
        # (persist 0) and with it (1, the default), and SSSP with near-far
        # (delta 400, 800, 1600).  Device time per run (best of 3) and per
        # iteration.
        hg = K.HostGraph.grid(2048, 2048, 201, wmin=1, wmax=100)
        dg = K.DeviceGraph(hg)
        dg.set_model(S.load_model())
        rows = []
        for algo in ("sssp", "bfs"):
            ref = None
            runs = [(sched, persist, 0) for sched in ("plain", "learned") for persist in (0, 1)]
            if algo == "sssp":
                runs += [("plain", 1, delta) for delta in (400, 800, 1600)]
            for sched, persist, delta in runs:
                best, info = None, None
                for _ in range(3):
                    dg.init(algo, 0)
                    info = dg.run(algo, sched=sched, persist=persist, delta=delta)
                    best = info["ms"] if best is None else min(best, info["ms"])
                vals = dg.values()
                ref = vals if ref is None else ref
                assert np.array_equal(vals, ref), f"grid {algo} {sched} {persist} differs"
                dg.init(algo, 0)
                dg.run(algo, sched=sched, persist=persist, delta=delta, keep_stats=True)
                nloop = sum(1 for x in dg.stats() if K.VARIANTS[x["variant"]] == "persist")
                it = info["iterations"]
                rows.append({"algo": algo, "schedule": sched, "persist": persist,
                             "delta": delta, "device_ms": round(best, 3), "iterations": it,
                             "loop_iterations": nloop,
                             "us_per_iteration": round(1e3 * best / max(1, it), 2)})
                log(f"highdiam {algo} {sched} persist={persist} delta={delta}: {best:.2f} ms, "
                    f"{it} it ({nloop} in the loop), {1e3 * best / max(1, it):.2f} us/it")
        dg.close()
        out["highdiam"] = {"workload": "SSSP and BFS from vertex 0 of the 2048 x 2048 "
                                       "4-neighbour grid, weights U[1,100], seed 201",
                           "runs": rows}

    if "graphs1000" in work:
        # 1,000 independent R-MAT graphs, scale U{10..16}, edge factor U{8..32},
        # one source each (max out-degree), one run of the learned schedule.
        tg = time.time()
        g, voff, gsrc = K.HostGraph.rmat_batch(1000, 1000, scale_lo=10, scale_hi=16, ef_lo=8,
                                               ef_hi=32, wmin=1, wmax=255)
        gR, gC, gE, goR, goC, goE = g.arrays()
        gout = np.diff(goR.astype(np.int64))
        log(f"graphs1000: n={len(gR) - 1} m={len(gC)} built in {time.time() - tg:.1f}s")
        dg = K.DeviceGraph(g)
        dg.set_model(S.load_model())
        res = {}
        tuned = {}
        for sched in ("plain", "learned"):
            tuned[sched], _, _ = dg.autotune("sssp", gsrc, sched=sched)
            log(f"graphs1000: tuned {sched} shape {tuned[sched]}")

            def gstep(sched=sched):
                dg.init_sources("sssp", gsrc)
                return dg.run("sssp", sched=sched, **tuned[sched])["ms"]
            wall, dev = time_steps(gstep, max(1, a.steps // 2), max(1, a.warmup // 2))
            res[sched] = (wall, dev)
        vals = dg.values()
        teps = int(gout[vals != K.INF].sum())
        out["graphs1000"] = {
            "workload": "SSSP on 1,000 R-MAT graphs (scale U{10..16}, edge factor U{8..32}, "
                        "seed 1000) in one run, one max-out-degree source per graph",
            "n": int(len(gR) - 1), "m": int(len(gC)), "reached_edges": teps,
            "tuned_shapes": tuned,
            "plain": {"ms_per_step": round(res["plain"][0], 3),
                      "gteps": round(teps / (res["plain"][0] * 1e-3) / 1e9, 3)},
            "learned": {"ms_per_step": round(res["learned"][0], 3),
                        "gteps": round(teps / (res["learned"][0] * 1e-3) / 1e9, 3)}}
        log(f"graphs1000: plain {res['plain'][0]:.2f} ms, learned {res['learned'][0]:.2f} ms")
    out["config"] = cfg
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
