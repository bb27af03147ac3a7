"""The learned iteration scheduler: data, training and the shipped model.

At every iteration the engine's last workgroup evaluates a small network
(10 features of the coming frontier -> 32 ReLU -> 8 outputs) and runs the
delta variant with the smallest output.  Output k is trained to predict
``min(log2(t_k / t_best), CLIP)``: how much slower variant k+1 (PUSH_T ..
PULL_LB) is than the fastest one from the same state, as timed by the
recorder (``DeviceGraph.record``) on seeded R-MAT graphs.  By default a
variant's time is its iteration plus the rest of the run with the forced
rollout variant (edge-balanced push), so a choice is also charged for the
frontier it leaves behind (one step of policy improvement over the rollout
policy); ``--rollout none`` times the iteration alone.

Training runs on the GPU through the library's HIP kernels
(``kites_mlp_train``: forward, backward and Adam); this module only prepares
arrays.  The reference has no counterpart: its schedule is one fixed kernel
per iteration (kites/device/compute.cuh:69-117).
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import check_mlp, lib
from .graph import DeviceGraph, HostGraph

MLP_IN, MLP_HID, MLP_OUT = _lib.MLP_IN, 32, 8
NPARAM = MLP_HID * MLP_IN + MLP_HID + MLP_OUT * MLP_HID + MLP_OUT
CLIP = 4.0
# Learned runs keep the plain choice unless the network predicts another
# variant faster by more than this many log2 units.
DEFAULT_MARGIN = 0.0
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models",
                          "scheduler.npz")

# Graphs the shipped model is trained on; evaluation graphs use other seeds.
TRAIN_SCALES = (16, 17, 18, 19, 20, 21)
TRAIN_SEEDS = (101, 102, 103)
EVAL_SEEDS = (201,)


def _f32(a, shape=None):
    a = np.ascontiguousarray(a, np.float32)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return a


def _p(a):
    return None if a is None else a.ctypes.data_as(_lib.f32p)


def init_weights(seed=0):
    """He-initialised weights in the flat layout W1 | b1 | W2 | b2."""
    rng = np.random.default_rng(seed)
    w1 = rng.normal(0, np.sqrt(2.0 / MLP_IN), (MLP_HID, MLP_IN))
    w2 = rng.normal(0, np.sqrt(1.0 / MLP_HID), (MLP_OUT, MLP_HID))
    return np.concatenate([w1.ravel(), np.zeros(MLP_HID), w2.ravel(),
                           np.zeros(MLP_OUT)]).astype(np.float32)


def forward(w, X, device=0):
    """Batch inference on the GPU: Y [B, 8]."""
    w, X = _f32(w, (NPARAM,)), _f32(X)
    B = X.shape[0]
    assert X.shape == (B, MLP_IN) and B > 0
    Y = np.empty((B, MLP_OUT), np.float32)
    check_mlp(lib().kites_mlp_forward(device, _p(w), _p(X), B, _p(Y)))
    return Y


def loss_grad(w, X, T, mask=None, device=0):
    """Loss (mean masked squared error) and its gradient, on the GPU."""
    w, X = _f32(w, (NPARAM,)), _f32(X)
    B = X.shape[0]
    T = _f32(T, (B, MLP_OUT))
    M = None if mask is None else _f32(mask, (B, MLP_OUT))
    g = np.empty(NPARAM, np.float32)
    loss = ctypes.c_float()
    check_mlp(lib().kites_mlp_loss_grad(device, _p(w), _p(X), _p(T), _p(M), B, _p(g),
                                        ctypes.byref(loss)))
    return float(loss.value), g


def train(w, X, T, mask=None, epochs=2000, lr=3e-3, device=0):
    """Full-batch Adam on the GPU; returns (weights, first loss, last loss)."""
    w = _f32(w, (NPARAM,)).copy()
    X = _f32(X)
    B = X.shape[0]
    T = _f32(T, (B, MLP_OUT))
    M = None if mask is None else _f32(mask, (B, MLP_OUT))
    l0, l1 = ctypes.c_float(), ctypes.c_float()
    check_mlp(lib().kites_mlp_train(device, _p(w), _p(X), _p(T), _p(M), B, epochs, lr,
                                    ctypes.byref(l0), ctypes.byref(l1)))
    return w, float(l0.value), float(l1.value)


def targets(ms):
    """Relative log2 slowdown of each variant against the fastest, clipped."""
    ms = np.maximum(np.asarray(ms, np.float64), 1e-4)
    t = np.log2(ms / ms.min(axis=1, keepdims=True))
    return np.minimum(t, CLIP).astype(np.float32)


TRAIN_EDGE_FACTORS = (16, 32)


def rmat_graph(scale, seed, edge_factor=16):
    return HostGraph.rmat(scale, edge_factor, seed, wmin=1, wmax=255)


def sources(host, k):
    """k sources with out-edges: the biggest hub, then evenly spaced ids."""
    R, C, E, oR, oC, oE = host.arrays()
    outdeg = np.diff(oR.astype(np.int64))
    cand = np.flatnonzero(outdeg > 0)
    picks = [int(np.argmax(outdeg))]
    for i in range(1, k):
        picks.append(int(cand[(i * len(cand)) // k]))
    return picks


def record_graph(dg, host, algos=("sssp", "bfs"), nsrc=2, log=None, **kw):
    """Recorded samples of one graph: list of (features, ms) arrays."""
    out = []
    for algo in algos:
        for src in sources(host, nsrc):
            dg.init(algo, src)
            feat, ms, it, pv = dg.record(algo, **kw)
            out.append((feat, ms))
            if log:
                log(f"  {algo} src {src}: {len(feat)} states, "
                    f"{float(ms.min(axis=1).sum()):.2f} ms best-variant total")
    return out


def train_graph(family, scale, seed, ef):
    """A training graph: R-MAT (rmat) or Erdos-Renyi with ef * n draws (er),
    2^scale vertices."""
    if family == "rmat":
        return rmat_graph(scale, seed, ef)
    if family == "er":
        n = 1 << scale
        return HostGraph.erdos_renyi(n, ef * n, seed)
    raise ValueError(f"unknown graph family {family}")


def record_dataset(scales=TRAIN_SCALES, seeds=TRAIN_SEEDS, log=print, device=0, nsrc=2,
                   edge_factors=TRAIN_EDGE_FACTORS, families=("rmat",), **kw):
    """SSSP and BFS from nsrc sources on each directed graph of each family,
    and min-label CC on its symmetrized copy."""
    X, Y = [], []
    for fam, scale in ((f, s) for f in families for s in scales):
        for seed, ef in ((s, e) for s in seeds for e in edge_factors):
            h = train_graph(fam, scale, seed, ef)
            dg = DeviceGraph(h, device)
            if log:
                log(f"{fam} scale {scale} edge factor {ef} seed {seed}: n {dg.n} m {dg.m}")
            for feat, ms in record_graph(dg, h, nsrc=nsrc, log=log, **kw):
                X.append(feat)
                Y.append(ms)
            dg.close()
            sym = h.symmetrize()
            dg = DeviceGraph(sym, device)
            for feat, ms in record_graph(dg, sym, algos=("cc",), nsrc=1, log=log, **kw):
                X.append(feat)
                Y.append(ms)
            dg.close()
    return np.concatenate(X), np.concatenate(Y)


def save_model(w, path=MODEL_PATH, **meta):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, weights=_f32(w, (NPARAM,)),
             **{k: np.asarray(v) for k, v in meta.items()})


def load_model(path=MODEL_PATH):
    """The shipped weights (numpy arrays only, no pickled objects)."""
    with np.load(path, allow_pickle=False) as z:
        return _f32(z["weights"], (NPARAM,))


def sample_weights(ms):
    """Per-state weights proportional to the state's best time: a wrong
    choice costs in proportion to how long the state takes, so the loss
    spends the network's capacity on the expensive (dense) states."""
    best = np.maximum(np.asarray(ms, np.float64).min(axis=1), 1e-4)
    w = np.clip(best / best.mean(), 0.05, 20.0)
    return w.astype(np.float32)


def settled_pull_times(X, ms):
    """A settled run (last feature 1) runs the thread-per-vertex pull
    whichever pull the network picks (the only pull that stops at the first
    updated in-neighbour), so every pull output of such a state is trained
    on that pull's time."""
    ms = np.array(ms, np.float32, copy=True)
    st = np.asarray(X)[:, MLP_IN - 1] > 0.5
    ms[st, 5:8] = ms[st, 4:5]
    return ms


def fit(X, ms, seed=0, epochs=4000, lr=3e-3, device=0, weighted=True):
    """Trains the network on recorded (features, ms) pairs.  The loss mask
    carries the square root of the state weights (the loss squares it)."""
    ms = settled_pull_times(X, ms)
    T = targets(ms)
    w0 = init_weights(seed)
    mask = None
    if weighted:
        mask = np.repeat(np.sqrt(sample_weights(ms))[:, None], MLP_OUT, axis=1)
    return train(w0, X, T, mask=mask, epochs=epochs, lr=lr, device=device)


def eval_graphs(scale=22):
    """The held-out evaluation graphs: (name, directed graph, symmetrized)."""
    n = 1 << scale
    side = 1 << (scale // 2)
    out = []
    for name, mk in (("rmat%d" % scale, lambda: rmat_graph(scale, EVAL_SEEDS[0])),
                     ("er%d" % scale, lambda: HostGraph.erdos_renyi(n, 16 * n, EVAL_SEEDS[0])),
                     ("grid%dx%d" % (side, side),
                      lambda: HostGraph.grid(side, side, EVAL_SEEDS[0], wmin=1, wmax=100))):
        h = mk()
        out.append((name, h))
    return out


ALPHA_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models",
                          "plain_alpha.json")
# Push/pull thresholds the alpha sweep tries (pull when out-edges * alpha > m);
# 1e9 pulls whenever the frontier has an out-edge.
ALPHA_GRID = (1.0, 2.0, 4.0, 8.0, 14.0, 24.0, 48.0, 96.0, 1e9)


def tuned_alpha(path=ALPHA_PATH):
    """The plain schedule's threshold tuned on the training graphs (the
    `alpha` command), or None before it has been run."""
    import json
    try:
        with open(path) as f:
            return float(json.load(f)["alpha"])
    except (OSError, KeyError, ValueError):
        return None


def baseline_schedules(alpha=None):
    """The schedules the learned one is compared with: the plain schedule at
    its default threshold (14), at the threshold tuned on the training
    graphs, and edge-balanced push every iteration."""
    out = {"plain": dict(sched="plain")}
    a = tuned_alpha() if alpha is None else alpha
    if a is not None:
        out["plain_tuned"] = dict(sched="plain", alpha=a)
    out["push_only"] = dict(sched="forced", forced_variant="push_lb")
    return out


def ab_compare(dg, host, algo, w, repeats=5, check=None, margin=DEFAULT_MARGIN, alpha=None):
    """Best-of-`repeats` device time of the learned schedule and of every
    baseline schedule from the same source; all results must be identical.
    `w` is one model's weights, or a dict {name: weights} of models each run
    as its own learned schedule (the first is "learned").  Returns a result
    dict with each time and learned's gain over each baseline."""
    R, C, E, oR, oC, oE = host.arrays()
    src = int(np.argmax(np.diff(oR.astype(np.int64))))
    models = w if isinstance(w, dict) else {"learned": w}
    scheds = dict(baseline_schedules(alpha))
    for name in models:
        scheds[name] = dict(sched="learned", margin=margin)
    best, vals, iters = {}, {}, {}
    for _ in range(repeats):
        for name, kw in scheds.items():
            if name in models:
                dg.set_model(models[name])
            v, info = dg.solve(algo, src, **kw)
            best[name] = min(best.get(name, 1e30), info["ms"])
            iters[name] = info["iterations"]
            if name in vals:
                assert np.array_equal(vals[name], v), f"{name} run not repeatable"
            vals[name] = v
    first = next(iter(models))
    for name in scheds:
        assert np.array_equal(vals[name], vals[first]), f"{name} result differs from {first}"
    if check is not None:
        check(vals["plain"], src)
    out = dict(algo=algo, source=src)
    for name in scheds:
        out[name + "_ms"] = best[name]
        out[name + "_iters"] = iters[name]
    for m in models:
        tag = "" if m == "learned" else "_" + m
        out["gain" + tag] = best["plain"] / best[m]
        for name in scheds:
            if name not in models and name != "plain":
                out["gain" + tag + "_vs_" + name] = best[name] / best[m]
    return out


def sweep_alpha(scales=TRAIN_SCALES, seeds=TRAIN_SEEDS, edge_factors=(16, 32), grid=ALPHA_GRID,
                repeats=3, log=print, device=0):
    """Tunes the plain schedule's push/pull threshold on the training graphs
    only: for every (graph, functor set, source) the best-of-`repeats` time
    at each alpha, normalised by the fastest alpha; the alpha with the least
    geometric mean of those ratios wins.  Returns (alpha, table)."""
    ratios = {a: [] for a in grid}
    for scale in scales:
        for seed in seeds:
            for ef in edge_factors:
                h = rmat_graph(scale, seed, ef)
                R, C, E, oR, oC, oE = h.arrays()
                outdeg = np.diff(oR.astype(np.int64))
                srcs = [int(np.argmax(outdeg)),
                        int(np.flatnonzero(outdeg > 0)[len(np.flatnonzero(outdeg > 0)) // 2])]
                dg = DeviceGraph(h, device)
                runs = [("sssp", s) for s in srcs] + [("bfs", s) for s in srcs]
                for algo, src in runs:
                    t = {}
                    ref = None
                    for a in grid:
                        for _ in range(repeats):
                            v, info = dg.solve(algo, src, sched="plain", alpha=a)
                            t[a] = min(t.get(a, 1e30), info["ms"])
                        if ref is None:
                            ref = v
                        assert np.array_equal(v, ref), f"alpha {a} differs"
                    tb = min(t.values())
                    for a in grid:
                        ratios[a].append(t[a] / tb)
                    log(f"rmat{scale} seed {seed} ef {ef} {algo} src {src}: " +
                        " ".join(f"{a:g}:{t[a]:.3f}" for a in grid))
                dg.close()
                sym = h.symmetrize()
                dg = DeviceGraph(sym, device)
                t = {}
                for a in grid:
                    for _ in range(repeats):
                        _, info = dg.solve("cc", 0, sched="plain", alpha=a)
                        t[a] = min(t.get(a, 1e30), info["ms"])
                tb = min(t.values())
                for a in grid:
                    ratios[a].append(t[a] / tb)
                log(f"rmat{scale} seed {seed} ef {ef} cc (symmetrized): " +
                    " ".join(f"{a:g}:{t[a]:.3f}" for a in grid))
                dg.close()
    table = {a: float(np.exp(np.mean(np.log(ratios[a])))) for a in grid}
    alpha = min(table, key=table.get)
    return alpha, table


MARGIN_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "models",
                           "learned_margin.json")
MARGIN_GRID = (0.0, 0.1, 0.2, 0.35, 0.5, 1.0)


def sweep_margin(w, scales=TRAIN_SCALES, seeds=TRAIN_SEEDS, edge_factors=(16, 32),
                 families=("rmat", "er"), grid=MARGIN_GRID, repeats=3, log=print, device=0):
    """Tunes the learned schedule's margin (keep the plain choice unless the
    network predicts another variant faster by more than this many log2
    units) on the training graphs only, as `sweep_alpha` tunes alpha: the
    margin with the least geometric mean of best-of-`repeats` time over the
    per-case fastest margin wins.  Returns (margin, table)."""
    ratios = {g: [] for g in grid}

    def one(dg, algo, src, name):
        t, ref = {}, None
        for g in grid:
            for _ in range(repeats):
                v, info = dg.solve(algo, src, sched="learned", margin=g)
                t[g] = min(t.get(g, 1e30), info["ms"])
            if ref is None:
                ref = v
            assert np.array_equal(v, ref), f"margin {g} differs"
        tb = min(t.values())
        for g in grid:
            ratios[g].append(t[g] / tb)
        log(f"{name} {algo} src {src}: " + " ".join(f"{g:g}:{t[g]:.3f}" for g in grid))

    for fam, scale in ((f, s) for f in families for s in scales):
        for seed, ef in ((s, e) for s in seeds for e in edge_factors):
            h = train_graph(fam, scale, seed, ef)
            dg = DeviceGraph(h, device)
            dg.set_model(w)
            name = f"{fam}{scale} seed {seed} ef {ef}"
            for src in sources(h, 2):
                for algo in ("sssp", "bfs"):
                    one(dg, algo, src, name)
            dg.close()
            sym = h.symmetrize()
            dg = DeviceGraph(sym, device)
            dg.set_model(w)
            one(dg, "cc", 0, name + " (symmetrized)")
            dg.close()
    table = {g: float(np.exp(np.mean(np.log(ratios[g])))) for g in grid}
    return min(table, key=table.get), table


def main(argv=None):
    import argparse
    import time
    ap = argparse.ArgumentParser(description="record timings / train the scheduler network")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("record")
    r.add_argument("out")
    r.add_argument("--scales", default=",".join(map(str, TRAIN_SCALES)))
    r.add_argument("--seeds", default=",".join(map(str, TRAIN_SEEDS)))
    r.add_argument("--nsrc", type=int, default=2)
    r.add_argument("--edge-factors", default=",".join(map(str, TRAIN_EDGE_FACTORS)))
    r.add_argument("--rollout", default="push_lb", help="forced variant after the fork, or none")
    r.add_argument("--families", default="rmat", help="training graph families: rmat,er")
    t = sub.add_parser("train")
    t.add_argument("data", nargs="+", help="recorded data files (concatenated)")
    t.add_argument("--out", default=MODEL_PATH)
    t.add_argument("--epochs", type=int, default=4000)
    t.add_argument("--lr", type=float, default=3e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--unweighted", action="store_true")
    e = sub.add_parser("ab")
    e.add_argument("--scale", type=int, default=22)
    e.add_argument("--repeats", type=int, default=5)
    e.add_argument("--model", default=MODEL_PATH,
                   help="model file, or a comma list name=file,... compared side by side")
    e.add_argument("--margin", type=float, default=DEFAULT_MARGIN)
    al = sub.add_parser("alpha", help="tune the plain schedule's threshold on the training graphs")
    al.add_argument("--scales", default=",".join(map(str, TRAIN_SCALES)))
    al.add_argument("--seeds", default=",".join(map(str, TRAIN_SEEDS)))
    al.add_argument("--out", default=ALPHA_PATH)
    mg = sub.add_parser("margin", help="tune the learned schedule's margin on the training graphs")
    mg.add_argument("--scales", default=",".join(map(str, TRAIN_SCALES)))
    mg.add_argument("--seeds", default=",".join(map(str, TRAIN_SEEDS)))
    mg.add_argument("--families", default="rmat,er")
    mg.add_argument("--model", default=MODEL_PATH)
    mg.add_argument("--out", default=MARGIN_PATH)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    if a.cmd == "record":
        t0 = time.time()
        X, ms = record_dataset([int(x) for x in a.scales.split(",")],
                               [int(x) for x in a.seeds.split(",")], log=log, nsrc=a.nsrc,
                               edge_factors=[int(x) for x in a.edge_factors.split(",")],
                               families=tuple(a.families.split(",")),
                               rollout=0 if a.rollout == "none" else a.rollout)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        np.savez(a.out, X=X, ms=ms)
        log(f"{len(X)} samples in {time.time() - t0:.1f} s -> {a.out}")
    elif a.cmd == "alpha":
        import json
        alpha, table = sweep_alpha([int(x) for x in a.scales.split(",")],
                                   [int(x) for x in a.seeds.split(",")], log=log)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with open(a.out, "w") as f:
            json.dump({"alpha": alpha, "geomean_ratio_to_best": {str(k): v for k, v in table.items()},
                       "graphs": f"R-MAT scales {a.scales}, seeds {a.seeds}, edge factors 16 and 32",
                       "functor_sets": "SSSP and BFS from 2 sources each, CC symmetrized"}, f,
                      indent=1)
        log(f"tuned alpha {alpha} (geometric mean ratio to best {table[alpha]:.3f}) -> {a.out}")
    elif a.cmd == "margin":
        import json
        margin, table = sweep_margin(load_model(a.model), [int(x) for x in a.scales.split(",")],
                                     [int(x) for x in a.seeds.split(",")],
                                     families=tuple(a.families.split(",")), log=log)
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        with open(a.out, "w") as f:
            json.dump({"margin": margin,
                       "geomean_ratio_to_best": {str(k): v for k, v in table.items()},
                       "graphs": f"{a.families} scales {a.scales}, seeds {a.seeds}, edge factors "
                                 "16 and 32",
                       "functor_sets": "SSSP and BFS from 2 sources each, CC symmetrized"}, f,
                      indent=1)
        log(f"tuned margin {margin} (geometric mean ratio to best {table[margin]:.3f}) -> {a.out}")
    elif a.cmd == "ab":
        import json
        if "=" in a.model:
            w = {k: load_model(v) for k, v in (x.split("=", 1) for x in a.model.split(","))}
        else:
            w = load_model(a.model)
        for name, h in eval_graphs(a.scale):
            dg = DeviceGraph(h)
            for algo in ("sssp", "bfs"):
                r = ab_compare(dg, h, algo, w, a.repeats, margin=a.margin)
                log(json.dumps(dict(graph=name, n=dg.n, m=dg.m, **r)))
            dg.close()
            sym = h.symmetrize()
            dg = DeviceGraph(sym)
            r = ab_compare(dg, sym, "cc", w, a.repeats, margin=a.margin)
            log(json.dumps(dict(graph=name + "-sym", n=dg.n, m=dg.m, **r)))
            dg.close()
    else:
        Xs, Ms = [], []
        for path in a.data:
            with np.load(path, allow_pickle=False) as z:
                Xs.append(z["X"])
                Ms.append(z["ms"])
        X, ms = np.concatenate(Xs), np.concatenate(Ms)
        w, l0, l1 = fit(X, ms, a.seed, a.epochs, a.lr, weighted=not a.unweighted)
        ms = settled_pull_times(X, ms)
        Y = forward(w, X)
        pick = Y.argmin(1)
        best = ms.min(1)
        regret = ms[np.arange(len(ms)), pick].sum() / best.sum()
        log(f"loss {l0:.4f} -> {l1:.4f}; chosen / best time on the training set {regret:.3f}")
        save_model(w, a.out, loss=l1, samples=len(X), epochs=a.epochs, lr=a.lr, seed=a.seed,
                   weighted=not a.unweighted)
        log(f"saved {a.out}")


if __name__ == "__main__":
    main()
