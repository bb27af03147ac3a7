"""The learned scheduler on the GPU: the HIP training kernels against the
float64 oracle (oracle/mlp_oracle.py), the recorder, and learned runs.

The scheduler never changes results: every learned run is compared bit for
bit with the oracle fixpoint / Dijkstra.  The device-side inference inside
the iteration kernel is checked by replaying the recorded statistics of a
learned run through the oracle network: each iteration's variant must be
the oracle's argmin wherever the two best predictions are not a near tie.
"""
import time

import numpy as np
import pytest

from conftest import case

pytestmark = pytest.mark.gpu

K = pytest.importorskip("kites_amd")
from kites_amd import scheduler as S  # noqa: E402
from oracle import liboracle as O  # noqa: E402
from oracle import mlp_oracle as M  # noqa: E402


def _problem(seed, B):
    rng = np.random.default_rng(seed)
    w = (rng.normal(0, 0.4, M.NPARAM)).astype(np.float32)
    X = rng.uniform(0, 1, (B, M.IN)).astype(np.float32)
    T = rng.uniform(0, 4, (B, M.OUT)).astype(np.float32)
    return w, X, T


def test_mlp_forward_matches_oracle():
    case("mlp forward B=5000")
    w, X, _ = _problem(1, 5000)
    Y = S.forward(w, X)
    Yo, _ = M.forward(w, X)
    assert np.allclose(Y, Yo, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("masked", [False, True])
def test_mlp_gradient_matches_oracle(masked):
    case(f"mlp loss_grad masked={masked}")
    w, X, T = _problem(2, 3000)
    mask = None
    if masked:
        mask = (np.random.default_rng(3).uniform(size=T.shape) > 0.25).astype(np.float32)
    loss, g = S.loss_grad(w, X, T, mask)
    lo, go = M.loss_grad(w, X, T, mask)
    assert loss == pytest.approx(lo, rel=1e-5)
    # fp32 sums of 3000 terms against float64: relative 1e-4 of the largest entry
    assert np.allclose(g, go, rtol=1e-4, atol=1e-4 * np.abs(go).max())


def test_mlp_training_follows_oracle_adam():
    case("mlp train 30 epochs vs oracle adam")
    w, X, T = _problem(4, 512)
    wg, l0, l1 = S.train(w, X, T, epochs=30, lr=1e-3)
    wo, lo0, lo1 = M.adam(w, X, T, epochs=30, lr=1e-3)
    assert l0 == pytest.approx(lo0, rel=1e-5)
    assert l1 == pytest.approx(lo1, rel=1e-3) and l1 < l0
    # steps are at most lr each; near-zero gradients may flip sign in fp32
    assert np.abs(wg - wo).max() < 30 * 1e-3 * 2
    assert np.mean(np.abs(wg - wo) < 1e-5) > 0.95


def test_mlp_training_fits():
    case("mlp train fits a learnable target")
    rng = np.random.default_rng(5)
    X = rng.uniform(0, 1, (2048, M.IN)).astype(np.float32)
    T = np.stack([X[:, k % M.IN] * (1 + k) for k in range(M.OUT)], 1).astype(np.float32)
    _, l0, l1 = S.train(S.init_weights(0), X, T, epochs=1500, lr=3e-3)
    assert l1 < 0.02 * l0


@pytest.mark.parametrize("rollout", [0, "push_lb"])
def test_recorder_leaves_result_exact(rollout):
    case(f"recorder sssp rmat14 rollout={rollout}")
    h = K.HostGraph.rmat(14, 16, 7, wmin=1, wmax=255)
    R, C, E, oR, oC, oE = h.arrays()
    src = int(np.argmax(np.diff(oR)))
    dg = K.DeviceGraph(h)
    dg.init("sssp", src)
    feat, ms, it, pv = dg.record("sssp", rollout=rollout)
    assert np.array_equal(dg.values(), O.dijkstra(R, C, E, src))
    assert len(feat) >= 3 and it[0] == 0 and np.all(np.diff(it.astype(np.int64)) == 1)
    assert np.all(ms > 0) and np.all(np.isfinite(ms))
    assert np.all((pv >= 1) & (pv <= 8))
    d = int(oR[src + 1] - oR[src])
    n, m = dg.n, dg.m
    assert np.allclose(feat[0], M.features(n, m, 1, d, d, 0), rtol=1e-6, atol=1e-7)
    # a normal run after recording starts from scratch and agrees
    vals, _ = dg.solve("sssp", src, sched="plain")
    assert np.array_equal(vals, O.dijkstra(R, C, E, src))


def _model():
    try:
        return S.load_model()
    except FileNotFoundError:
        pytest.skip("no trained scheduler model in kites_amd/models")


def _plain_variant(sdeg, m, alpha=14.0, settled=False):
    # plain schedule, edge-balanced granularity: pull when out-edges * alpha > m;
    # a settled run (BFS from a unit start) pulls thread per vertex
    if sdeg * alpha > m:
        return 5 if settled else 8
    return 4


def _quarter_octave(x):
    y = int(x) + 1
    lb = y.bit_length() - 1
    frac = ((y >> (lb - 2)) if lb >= 2 else (y << (2 - lb))) & 3
    return 4 * lb + frac


def _learned_key(count, sdeg, maxdeg, it):
    return ((_quarter_octave(count) & 0xff) << 24 | (_quarter_octave(sdeg) & 0xff) << 16 |
            ((_quarter_octave(maxdeg) >> 2) & 0xff) << 8 | (_quarter_octave(it) & 0xff))


SMALL_WORK = 1 << 18   # small-work guard: queue entries + out-edges at most this


def _guard(v, fb, sdeg, m, settled=False):
    # pull cost guard: the learned schedule pulls only if out-edges * 8 >= m
    # (settled runs: * 32); a settled run's pull is thread per vertex
    if 5 <= v <= 8 and sdeg * (32 if settled else 8) < m:
        return fb
    return 5 if (5 <= v <= 8 and settled) else v


@pytest.mark.parametrize("scale,algo,margin,min_inferred",
                         [(16, "sssp", 0.0, 1), (18, "bfs", 0.0, 0), (19, "sssp", 0.0, 1),
                          (19, "sssp", 0.15, 1)])
def test_learned_runs_exact_and_infer_like_oracle(scale, algo, margin, min_inferred):
    case(f"learned {algo} rmat{scale} seed 201 margin {margin}")
    w = _model()
    h = S.rmat_graph(scale, 201)
    R, C, E, oR, oC, oE = h.arrays()
    src = int(np.argmax(np.diff(oR)))
    dg = K.DeviceGraph(h)
    dg.set_model(w)
    vals, info = dg.solve(algo, src, sched="learned", keep_stats=True, margin=margin)
    ref = O.dijkstra(R, C, E, src) if algo == "sssp" else O.bfs(R, C, src)
    assert np.array_equal(vals, ref)
    st = dg.stats()
    assert len(st) == info["iterations"]
    # Replay the engine's decisions with the oracle network, including its
    # small-work guard (plain choice, no inference, cache untouched), its
    # decision cache (quantized features) and the pull guard.  Where fp32
    # and float64 predictions are within 1e-3 of a tie, every tied choice
    # is accepted.
    d = int(oR[src + 1] - oR[src])
    states = [(1, d, d)] + [(p["count"], p["sum_outdeg"], p["max_outdeg"]) for p in st[:-1]]
    cache_key, cache_allowed = None, None
    inferred = guarded = 0
    settled = algo == "bfs"   # BFS from one source at 0: levels settle
    for i, (cnt, sdeg, mdeg) in enumerate(states):
        fb = _plain_variant(sdeg, dg.m, settled=settled)
        if settled:
            # settled runs take the plain rule, no inference
            assert st[i]["variant"] == fb, i
            guarded += 1
            continue
        if cnt + sdeg <= SMALL_WORK:
            assert st[i]["variant"] == fb, i
            guarded += 1
            continue
        key = _learned_key(cnt, sdeg, mdeg, i)
        if key != cache_key:
            x = M.features(dg.n, dg.m, cnt, sdeg, mdeg, i, settled)
            y = M.forward(w, x[None, :])[0][0]
            best = {int(k) + 1 for k in np.flatnonzero(y - y.min() <= 1e-3)}
            gap = y[fb - 1] - y.min() - margin
            allowed = {fb} if gap < -1e-3 else best if gap > 1e-3 else best | {fb}
            assert M.choose(w, x, fb, margin) in allowed
            cache_key, cache_allowed = key, allowed
        assert st[i]["variant"] in {_guard(v, fb, sdeg, dg.m, settled) for v in cache_allowed}, i
        inferred += 1
    # BFS frontiers are mostly small: its decisions may all be guarded
    assert inferred >= min_inferred and inferred + guarded == len(states)


def test_learned_against_plain():
    """Held-out graph (R-MAT-20, a seed never used for training): what the
    trained network gains over the plain schedule at its default threshold,
    at the threshold tuned on the training graphs, and over edge-balanced
    push in every iteration, for SSSP, BFS and CC (symmetrized), best of 8
    interleaved runs each, printed.  All results are identical.  Learned
    must be at least 0.95x as fast as push-only over the three functor sets
    (geometric mean) and within run-to-run noise of it on each (0.85x: these
    runs take 0.4-1.2 ms); the full A/B on the evaluation graphs is
    `python -m kites_amd.scheduler ab`."""
    w = _model()
    base = S.rmat_graph(20, 201)
    push_gains, plain_gains = [], []
    for algo in ("sssp", "bfs", "cc"):
        case(f"learned vs baselines {algo} rmat20 seed 201")
        h = base.symmetrize() if algo == "cc" else base
        R, C, E, oR, oC, oE = h.arrays()
        src = int(np.argmax(np.diff(oR)))
        dg = K.DeviceGraph(h)
        dg.set_model(w)
        ref = (O.dijkstra(R, C, E, src) if algo == "sssp" else O.bfs(R, C, src)
               if algo == "bfs" else O.cc_min_label(R, C))
        scheds = dict(S.baseline_schedules())
        scheds["learned"] = dict(sched="learned")
        best = {}
        for _ in range(8):
            for name, kw in scheds.items():
                vals, info = dg.solve(algo, src, **kw)
                assert np.array_equal(vals, ref), name
                best[name] = min(best.get(name, 1e9), info["ms"])
        gains = {k: best[k] / best["learned"] for k in best if k != "learned"}
        print(f"{algo}: " + ", ".join(f"{k} {best[k]:.3f} ms" for k in best) + "; learned gain " +
              ", ".join(f"{k} {g:.2f}x" for k, g in gains.items()), flush=True)
        push_gains.append(gains["push_only"])
        plain_gains.append(gains["plain"])
        dg.close()
    geo = float(np.exp(np.mean(np.log(push_gains))))
    print(f"learned over push-only: geometric mean {geo:.3f}", flush=True)
    assert geo >= 0.95 and min(push_gains) >= 0.85
    assert min(plain_gains) > 0.75


def test_learned_with_split_pull():
    """The learned schedule with the source-split pull gives the plain
    schedule's values bit for bit."""
    from kites_amd import scheduler as S
    h = K.HostGraph.rmat(15, 16, 61, wmin=1, wmax=255)
    d = K.DeviceGraph(h)
    d.set_model(S.load_model())
    _, _, _, oR, _, _ = h.arrays()
    src = int(np.argmax(np.diff(oR)))
    for algo in ("sssp", "bfs"):
        ref, _ = d.solve(algo, src, sched="plain")
        got, _ = d.solve(algo, src, sched="learned", split_pull=True)
        assert np.array_equal(got, ref), algo


def test_learned_on_grid_equals_oracle():
    """High-diameter grid (thousands of small push iterations, where the
    learned schedule skips its degree statistics because the next decision
    is surely guarded): values equal the oracle's, BFS levels the closed
    form |dr| + |dc|."""
    rows, cols = 96, 80
    h = K.HostGraph.grid(rows, cols, 5, wmin=1, wmax=100)
    d = K.DeviceGraph(h)
    d.set_model(_model())
    R, C, E, *_ = h.arrays()
    case("learned on grid 96x80")
    sp, info = d.solve("sssp", 0, sched="learned")
    assert np.array_equal(sp, O.dijkstra(R, C, E, 0)) and info["iterations"] > rows
    lv, _ = d.solve("bfs", 0, sched="learned")
    rr, cc = np.divmod(np.arange(rows * cols), cols)
    assert np.array_equal(lv[:rows * cols], (rr + cc).astype(np.uint32))
