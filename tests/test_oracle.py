"""Pins the CPU oracle (oracle/kites_oracle.c) before it is trusted.

The oracle's Gauss-Seidel fixpoint restates the reference CPU path
(kites/graph/sanity_check.cuh:21-47).  Each functor set is checked against
an independent textbook algorithm, exhaustive Bellman-Ford on tiny graphs
and closed forms, so that a dropped term, a wrong comparison or a transposed
index fails here.  The reference ships no golden outputs; the Wiki-Vote
vector committed under tests/golden was written only after the oracle and
Dijkstra agreed on it, and is re-derived here.
"""
import gzip
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from oracle import graph_raw_oracle as gro
from oracle import liboracle as O

INF = O.INF


def random_csr(n, m, seed, wmax=20, symmetric=False, allow_dups=True):
    rng = np.random.default_rng(seed)
    s = rng.integers(0, n, m)
    d = rng.integers(0, n, m)
    keep = s != d
    s, d = s[keep], d[keep]
    if symmetric:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
    if not allow_dups:
        key = np.unique(s.astype(np.int64) * n + d)
        s, d = key // n, key % n
    w = rng.integers(0, wmax + 1, len(s)).astype(np.uint32)
    order = np.lexsort((s, d))
    s, d, w = s[order], d[order], w[order]
    R = np.zeros(n + 1, np.int64)
    np.add.at(R, d + 1, 1)
    return np.cumsum(R).astype(np.uint32), s.astype(np.uint32), w


def bellman_ford_brute(R, C, E, src, n):
    """Exhaustive relaxation, n rounds over every edge (pure Python)."""
    dist = [INF] * n
    dist[src] = 0
    for _ in range(n):
        for v in range(n):
            for e in range(R[v], R[v + 1]):
                u = C[e]
                if dist[u] != INF:
                    nd = min(dist[u] + int(E[e]), INF)
                    if nd < dist[v]:
                        dist[v] = nd
    return np.array(dist, np.uint32)


@pytest.mark.parametrize("seed", range(6))
def test_sssp_fixpoint_equals_dijkstra_and_bruteforce(seed):
    n = 12 if seed < 3 else 300
    R, C, E = random_csr(n, 4 * n, seed)
    src = seed % n
    V0, _ = O.sssp_init(n, src, R, C)
    V, sweeps = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
    assert sweeps >= 1
    assert np.array_equal(V, O.dijkstra(R, C, E, src))
    if n <= 12:
        assert np.array_equal(V, bellman_ford_brute(R, C, E, src, n))


@pytest.mark.parametrize("seed", range(4))
def test_bfs_and_widest(seed):
    n = 400
    R, C, E = random_csr(n, 3 * n, 100 + seed, wmax=50)
    src = 3
    V0, _ = O.sssp_init(n, src, R, C)
    Vb, _ = O.fixpoint(R, C, None, V0, O.BFS)
    assert np.array_equal(Vb, O.bfs(R, C, src))
    W0, _ = O.widest_init(n, src, R, C)
    Vw, _ = O.fixpoint(R, C, E, W0, O.WIDEST)
    assert np.array_equal(Vw, O.widest(R, C, E, src))


@pytest.mark.parametrize("seed", range(4))
def test_cc_equals_union_find(seed):
    n = 500
    R, C, E = random_csr(n, n // 2, 200 + seed, symmetric=True)
    V0, _ = O.cc_init(n)
    V, _ = O.fixpoint(R, C, E, V0, O.CC)
    lab = O.cc_min_label(R, C)
    assert np.array_equal(V, lab)
    # label = min id of the component: each label is its own label
    assert np.all(lab[lab] == lab) and np.all(lab <= np.arange(n))


def grid_csr(rows, cols, w=1):
    src, dst = [], []
    for r in range(rows):
        for c in range(cols):
            for dr, dc in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                r2, c2 = r + dr, c + dc
                if 0 <= r2 < rows and 0 <= c2 < cols:
                    src.append(r2 * cols + c2)
                    dst.append(r * cols + c)
    order = np.lexsort((src, dst))
    s, d = np.array(src)[order], np.array(dst)[order]
    R = np.zeros(rows * cols + 1, np.int64)
    np.add.at(R, d + 1, 1)
    return np.cumsum(R).astype(np.uint32), s.astype(np.uint32), np.full(len(s), w, np.uint32)


def test_closed_forms():
    # unit-weight grid: |dr| + |dc| (SSSP and BFS)
    R, C, E = grid_csr(9, 7)
    V0, _ = O.sssp_init(63, 0, R, C)
    rr, cc = np.divmod(np.arange(63), 7)
    for algo in (O.SSSP_SAT, O.BFS):
        V, _ = O.fixpoint(R, C, E, V0, algo)
        assert np.array_equal(V, (rr + cc).astype(np.uint32))
    # weighted path 0 -> 1 -> ... -> n-1: prefix sums
    n = 50
    w = np.arange(1, n, dtype=np.uint32)
    R = np.concatenate([[0], np.arange(n)]).astype(np.uint32)   # vertex v>0 has one in-edge
    C = np.arange(n - 1, dtype=np.uint32)
    V0, _ = O.sssp_init(n, 0, R, C)
    V, sweeps = O.fixpoint(R, C, w, V0, O.SSSP_SAT)
    assert np.array_equal(V, np.concatenate([[0], np.cumsum(w)]).astype(np.uint32))
    assert sweeps == 2   # in-order Gauss-Seidel settles a forward path in one sweep
    # star: every leaf at its edge weight
    n = 30
    R = np.concatenate([[0, 0], np.arange(1, n)]).astype(np.uint32)
    C = np.zeros(n - 1, np.uint32)
    E = (np.arange(n - 1) * 3 + 2).astype(np.uint32)
    V0, _ = O.sssp_init(n, 0, R, C)
    V, _ = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
    assert np.array_equal(V[1:], E)


def test_saturation_at_inf():
    # 0 -> 1 (w = 2^32 - 10) -> 2 (w = 100): the path sum overflows -> INF
    R = np.array([0, 0, 1, 2], np.uint32)
    C = np.array([0, 1], np.uint32)
    E = np.array([0xFFFFFFF6, 100], np.uint32)
    V0, _ = O.sssp_init(3, 0, R, C)
    V, _ = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
    assert V.tolist() == [0, 0xFFFFFFF6, INF]
    assert np.array_equal(V, O.dijkstra(R, C, E, 0))
    # the wrapping functor (reference example) does wrap: a distinct result
    V2, _ = O.fixpoint(R, C, E, V0, O.SSSP_WRAP)
    assert V2[2] == 90


def test_activation_rule_and_multicore_agree():
    R, C, E = random_csr(2000, 16000, 9, wmax=255, allow_dups=False)
    V0, F0 = O.sssp_init(2000, 17, R, C)
    gs, _ = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
    fr, rounds = O.frontier_rounds(R, C, E, V0, F0, O.SSSP_SAT)
    om, _, used = O.frontier_omp(R, C, E, V0, F0, O.SSSP_SAT, threads=4)
    assert np.array_equal(gs, fr) and np.array_equal(gs, om)
    assert used == 4 and rounds >= 2


def test_wikivote_golden_vector():
    raw = gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read().decode()
    g = gro.build(raw, gro.example_callback(30), 1, True, True, True)
    golden = np.load(os.path.join(GOLDEN, "wikivote_sssp_src30.npz"))["V"]
    V, _ = O.fixpoint(g["R"], g["C"], g["E"], g["V"], O.SSSP_WRAP)
    assert np.array_equal(V, golden)
    d = O.dijkstra(g["R"], g["C"], g["E"], 30)
    reach = d != INF
    assert np.array_equal(golden[reach], np.minimum(d[reach], g["V"][reach]))
    assert golden[30] == 0 and int(reach.sum()) == 2316


def test_c1_digest():
    K = pytest.importorskip("kites_amd")
    dig = json.load(open(os.path.join(GOLDEN, "c1_digest.json")))
    p = dig["params"]
    h = K.HostGraph.rmat(p["scale"], 0, p["seed"], wmin=p["wmin"], wmax=p["wmax"],
                         n_limit=p["n_limit"], exact_edges=p["exact_edges"])
    R, C, E, oR, oC, oE = h.arrays()
    sha = lambda *a: hashlib.sha256(b"".join(np.ascontiguousarray(x).tobytes() for x in a)).hexdigest()
    assert len(C) == 103689 and h.info()[1] == 7115
    assert sha(R, C, E) == dig["csr_sha256"]
    src = int(np.argmax(np.diff(oR)))
    assert src == dig["source"]
    V0, _ = O.sssp_init(len(R) - 1, src, R, C)
    V, _ = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
    assert sha(V) == dig["sssp_sha256"]


# ---- ordered (Jacobi) iterations with the defined combine order ------------

def _csr_from_edges(n, src, dst):
    src, dst = np.asarray(src, np.int64), np.asarray(dst, np.int64)
    o = np.lexsort((src, dst))
    R = np.zeros(n + 1, np.uint32)
    np.add.at(R, dst + 1, 1)
    R = np.cumsum(R).astype(np.uint32)
    oo = np.lexsort((dst, src))
    oR = np.zeros(n + 1, np.uint32)
    np.add.at(oR, src + 1, 1)
    oR = np.cumsum(oR).astype(np.uint32)
    return R, src[o].astype(np.uint32), o, oR, dst[oo].astype(np.uint32)


def test_ordered_tree_order_closed_form():
    """Four in-edges with contributions 1e8, 1, -1e8, 1 (float32): the
    documented tree gives (1e8 + 1) + (-1e8 + 1) = 1e8 + -1e8 = 0, where a
    left fold would give 1 and the grouping (1e8 - 1e8) + (1 + 1) would give 2."""
    R, C, _, oR, oC = _csr_from_edges(5, [1, 2, 3, 4], [0, 0, 0, 0])
    V0 = np.array([5, 1e8, 1, -1e8, 1], np.float32)
    act = np.array([1, 0, 0, 0, 0], np.uint8)
    V, it = O.ordered_pagerank(R, C, np.ones(4, np.float32), oR, oC, V0, act, base=0.0,
                               damp=1.0, tol=0.5)
    assert V[0] == 0.0 and it == 2 and np.array_equal(V[1:], V0[1:])


def test_ordered_exact_arithmetic_matches_float64():
    """A random DAG with unit weights, damping 1 and base 1: every value is
    an integer below 2^24, so float32 is exact and the ordered result must
    equal a float64 Jacobi sweep run to its fixpoint (unique on a DAG);
    in-lists longer than 64 exercise the chunking."""
    rng = np.random.default_rng(3)
    n, m = 400, 6000
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    keep = a < b                                  # edges go up: a DAG
    a, b = a[keep], b[keep]
    a = np.concatenate([a, np.arange(10, 90)])    # vertex 399 gets > 64 in-edges
    b = np.concatenate([b, np.full(80, 399)])
    R, C, order, oR, oC = _csr_from_edges(n, a, b)
    assert np.diff(R.astype(np.int64)).max() > 64
    V, it = O.ordered_pagerank(R, C, np.ones(len(C), np.float32), oR, oC,
                               np.zeros(n, np.float32), np.ones(n, np.uint8), base=1.0,
                               damp=1.0, tol=0.5)
    ref = np.zeros(n)
    while True:
        nxt = np.ones(n)
        np.add.at(nxt, b, ref[a])
        if np.array_equal(nxt, ref):
            break
        ref = nxt
    assert it > 2 and ref.max() < 2 ** 24
    assert np.array_equal(V.astype(np.float64), ref)


def _out_csr(R, C, E, n):
    """Out-edge CSR of an in-edge CSR (destinations sorted per source)."""
    d = np.repeat(np.arange(n, dtype=np.uint32), np.diff(R.astype(np.int64)))
    order = np.lexsort((d, C))
    oR = np.zeros(n + 1, np.int64)
    np.add.at(oR, C.astype(np.int64) + 1, 1)
    return np.cumsum(oR).astype(np.uint32), d[order], E[order]


def test_multicore_over_prepared_out_csr():
    R, C, E = random_csr(3000, 24000, 19, wmax=255, allow_dups=False)
    oR, oC, oE = _out_csr(R, C, E, 3000)
    for algo, w in ((O.SSSP_SAT, E), (O.BFS, None)):
        V0, F0 = O.sssp_init(3000, 5, R, C)
        gs, _ = O.fixpoint(R, C, w, V0, algo)
        om, _, used = O.frontier_omp_csr(R, C, w, oR, oC, None if w is None else oE, V0, F0,
                                         algo, threads=3)
        assert np.array_equal(gs, om) and used == 3


def test_check_cols_accepts_dijkstra_and_rejects_planted_mistakes():
    # weights >= 1: the certificate is exact for positive weights
    R, C, E = random_csr(1500, 9000, 23, wmax=40, allow_dups=False)
    E = np.maximum(E, 1).astype(np.uint32)
    src = np.array([0, 7, 99, 1499, 640], np.uint32)
    D = np.stack([O.dijkstra(R, C, E, int(s)) for s in src], 1)
    L = np.stack([O.bfs(R, C, int(s)) for s in src], 1)
    assert O.check_cols(R, C, E, D, src, threads=4) == (0, [0, 0, 0])
    assert O.check_cols(R, C, None, L, src, threads=4)[0] == 0
    reached = np.argwhere((D != INF) & (D != 0))
    rng = np.random.default_rng(1)
    for trial in range(12):
        v, k = reached[rng.integers(len(reached))]
        for delta in (+1, -1):
            bad = D.copy()
            bad[v, k] = np.uint32(int(bad[v, k]) + delta)
            assert O.check_cols(R, C, E, bad, src)[0] > 0, (v, k, delta)
        bad = D.copy()
        bad[v, k] = INF                      # reached vertex claimed unreached
        assert O.check_cols(R, C, E, bad, src)[0] > 0
    bad = D.copy()
    bad[src[2], 2] = 1                       # source not at zero
    assert O.check_cols(R, C, E, bad, src)[1][0] == 1
    # a column from the wrong source fails
    assert O.check_cols(R, C, E, D[:, [1, 0, 2, 3, 4]], src)[0] > 0
    # BFS levels are not SSSP distances of this weighting, and vice versa
    if np.any(E > 1):
        assert O.check_cols(R, C, E, L, src)[0] > 0


def test_dijkstra_ranges_equals_per_graph_dijkstra():
    rng = np.random.default_rng(4)
    sizes = [130, 1, 64, 500, 257]
    voff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    n = int(voff[-1])
    parts_s, parts_d, parts_w = [], [], []
    for i, sz in enumerate(sizes):
        m = 6 * sz
        s = rng.integers(0, sz, m) + voff[i]
        d = rng.integers(0, sz, m) + voff[i]
        parts_s.append(s)
        parts_d.append(d)
        parts_w.append(rng.integers(1, 256, m))
    s, d = np.concatenate(parts_s), np.concatenate(parts_d)
    w = np.concatenate(parts_w).astype(np.uint32)
    order = np.lexsort((s, d))
    s, d, w = s[order], d[order], w[order]
    R = np.zeros(n + 1, np.int64)
    np.add.at(R, d + 1, 1)
    R = np.cumsum(R).astype(np.uint32)
    C = s.astype(np.uint32)
    srcs = np.array([voff[i] + rng.integers(0, sz) for i, sz in enumerate(sizes)], np.uint32)
    got = O.dijkstra_ranges(R, C, w, voff, srcs, threads=3)
    for i in range(len(sizes)):
        lo, hi = int(voff[i]), int(voff[i + 1])
        lR = (R[lo:hi + 1] - R[lo]).astype(np.uint32)
        lC = (C[R[lo]:R[hi]] - lo).astype(np.uint32)
        exp = O.dijkstra(lR, lC, w[R[lo]:R[hi]], int(srcs[i]) - lo)
        assert np.array_equal(got[lo:hi], exp)
    with pytest.raises(ValueError):          # a source outside its graph
        O.dijkstra_ranges(R, C, w, voff, srcs[[1, 0, 2, 3, 4]])
