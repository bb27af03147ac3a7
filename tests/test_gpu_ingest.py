"""GPU graph build (kites_graph_from_edges) against the oracle and the host build.

The arrays must equal, bit for bit, what the oracle's restatement of the
reference parser and builder produces (kites/graph/graph_raw.h:58-190,
graph_csr.h:45-110; pinned in test_oracle.py on the Wiki-Vote fixture): the
in-CSR R, C, E and the out-lists dir_R / dir_C (vertex-per-group 1); the
weighted out-CSR must equal the host build's, and runs on the GPU-built
graph must equal Dijkstra.
"""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN, case

pytestmark = pytest.mark.gpu

K = pytest.importorskip("kites_amd")
from oracle import graph_raw_oracle as gro  # noqa: E402
from oracle import liboracle as O  # noqa: E402


def edge_list_of(text):
    """(src, dst) in file order, as the oracle's parser reads them."""
    src, dst = [], []
    for raw in text.split("\n"):
        if not raw or not ("0" <= raw[0] <= "9"):
            continue
        toks = gro._split(raw[:255])
        if len(toks) < 2:
            continue
        src.append(gro._atoll(toks[0]) & 0xFFFFFFFF)
        dst.append(gro._atoll(toks[1]) & 0xFFFFFFFF)
    return np.array(src, np.uint32), np.array(dst, np.uint32)


def text_of(src, dst, w):
    return "".join(f"{a} {b} {c}\n" for a, b, c in zip(src, dst, w))


def check_against_oracle(g, src, dst, w):
    text = text_of(src, dst, w)
    ref = gro.build(text, gro.example_callback(0), 1, True, True, True)
    R, C, E, oR, oC, oE, nzi, nzR = g.arrays()
    assert g.n == len(ref["R"]) - 1
    assert np.array_equal(R, ref["R"]) and np.array_equal(C, ref["C"])
    assert np.array_equal(E, ref["E"])
    assert np.array_equal(oR, ref["dir_R"]) and np.array_equal(oC, ref["dir_C"])
    nz = np.flatnonzero(np.diff(R.astype(np.int64)) > 0).astype(np.uint32)
    assert np.array_equal(nzi, nz) and np.array_equal(nzR, np.append(R[nz], R[-1]))
    return R, C, E, oR, oC, oE


def test_wikivote_fixture(tmp_path):
    raw = gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read().decode()
    src, dst = edge_list_of(raw)
    # weights from the reference example's callback (no third token)
    w = ((src.astype(np.uint64) + 1) * (dst.astype(np.uint64) + 2) % 127).astype(np.uint32)
    n = int(max(src.max(), dst.max())) + 1
    case("ingest Wiki-Vote fixture")
    g, ms = K.DeviceGraph.from_edges(n, src, dst, w)
    assert g.n == 8320 and g.m == 103689 and ms > 0
    ref = gro.build(raw, gro.example_callback(30), 1, True, True, True)
    R, C, E, oR, oC, oE, _, _ = g.arrays()
    assert np.array_equal(R, ref["R"]) and np.array_equal(C, ref["C"])
    assert np.array_equal(E, ref["E"])
    assert np.array_equal(oR, ref["dir_R"]) and np.array_equal(oC, ref["dir_C"])
    path = tmp_path / "wikivote.txt"
    path.write_text(raw)
    assert np.array_equal(oE, K.HostGraph.load_edgelist(path).arrays()[5])


@pytest.mark.parametrize("n,m,seed", [(300, 3000, 1), (1, 5, 2), (1000, 1, 4), (5000, 40000, 5)])
def test_random_lists_match_oracle(n, m, seed):
    """Duplicates, self-loops, isolated and edgeless vertices, vertex counts
    below and above a multiple of 128.  The vertex count is the largest id
    + 1, as the reference's parser derives it (graph_raw.h:90-94)."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    if m >= 10:
        src[: m // 10] = src[0]           # a hub with repeated edges
    w = rng.integers(0, 1000, m).astype(np.uint32)
    nid = int(max(src.max(), dst.max())) + 1
    case(f"ingest random n={n} m={m} seed={seed}")
    g, _ = K.DeviceGraph.from_edges(nid, src, dst, w)
    R, C, E, oR, oC, oE = check_against_oracle(g, src, dst, w)
    h = K.HostGraph.from_edges(nid, src, dst, w)
    assert np.array_equal(oE, h.arrays()[5])


def test_edgeless_graph():
    case("ingest without edges")
    g, _ = K.DeviceGraph.from_edges(129, [], [], [])
    R, C, E, oR, oC, oE, nzi, nzR = g.arrays()
    assert g.n == 256 and g.m == 0 and len(nzi) == 0
    assert not R.any() and not oR.any() and np.array_equal(nzR, [0])
    v, _ = g.solve("sssp", 5, sched="plain")
    assert v[5] == 0 and (np.delete(v, 5) == 0xFFFFFFFF).all()
    case("ingest without edges, then both multi-source paths")
    srcs = np.array([5, 0, 255, 5], np.uint32)
    exp = np.full((g.n, len(srcs)), 0xFFFFFFFF, np.uint32)
    exp[srcs, np.arange(len(srcs))] = 0
    vals, _ = g.batch("sssp", srcs)
    assert np.array_equal(vals, exp)
    lev, _ = g.bfs256(srcs)
    assert np.array_equal(lev, exp)


@pytest.mark.parametrize("dedup", [False, True])
def test_large_list_matches_host_build(dedup):
    """An R-MAT-shaped list of 4M draws (duplicates and self-loops) over 2^18
    vertices: the GPU build equals the host build, with and without dedup."""
    rng = np.random.default_rng(77)
    n, m = 1 << 18, 4_000_000
    # skewed endpoints: squares of uniforms concentrate on low ids
    src = (rng.random(m) ** 3 * n).astype(np.uint32)
    dst = (rng.random(m) ** 2 * n).astype(np.uint32)
    w = rng.integers(1, 256, m).astype(np.uint32)
    case(f"ingest large dedup={dedup}")
    g, ms = K.DeviceGraph.from_edges(n, src, dst, w, dedup=dedup)
    h = K.HostGraph.from_edges(n, src, dst, w, dedup=dedup)
    got = g.arrays()[:6]
    exp = h.arrays()
    for a, b in zip(got, exp):
        assert np.array_equal(a, b)
    if dedup:
        assert g.m < m
    else:
        assert g.m == m


def test_runs_on_gpu_built_graph():
    rng = np.random.default_rng(9)
    n, m = 20000, 300000
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    w = rng.integers(1, 100, m).astype(np.uint32)
    case("ingest then SSSP")
    g, _ = K.DeviceGraph.from_edges(n, src, dst, w, dedup=True)
    R, C, E = g.arrays()[:3]
    s = int(np.argmax(np.bincount(src, minlength=g.n)))
    exp = O.dijkstra(R, C, E, s)
    for kw in (dict(sched="plain"), dict(sched="forced", forced_variant="pull_lb"),
               dict(sched="forced", forced_variant="push_lb"), dict(sched="classic")):
        v, _ = g.solve("sssp", s, **kw)
        assert np.array_equal(v, exp), kw
    # the split pull, and the ordered mode (float PageRank) on the same graph:
    # identical to the host-built graph's, and to the oracle's PageRank
    v, _ = g.solve("sssp", s, sched="plain", split_pull=True)
    assert np.array_equal(v, exp)
    # both multi-source paths on the GPU-built in-lists (nzi, nzR, chunk starts)
    case("ingest then 64-source batch and 100-source bit-parallel BFS")
    srcs = np.concatenate([[s], rng.choice(np.delete(np.arange(n), s), 99, replace=False)])
    srcs = srcs.astype(np.uint32)
    vals, _ = g.batch("sssp", srcs[:64])
    assert np.array_equal(vals, np.stack([O.dijkstra(R, C, E, int(x)) for x in srcs[:64]], 1))
    lev, _ = g.bfs256(srcs)
    assert np.array_equal(lev, np.stack([O.bfs(R, C, int(x)) for x in srcs], 1))
    case("ingest then PageRank")
    h = K.HostGraph.from_edges(n, src, dst, w, dedup=True)
    hg = K.DeviceGraph(h)
    pr_g, ig = g.pagerank()
    pr_h, ih = hg.pagerank()
    assert ig["iterations"] == ih["iterations"]
    assert np.array_equal(pr_g.view(np.uint32), pr_h.view(np.uint32))
    R, C, _, oR, oC, _ = g.arrays()[:6]
    W = np.float32(1.0) / np.diff(oR.astype(np.int64)).astype(np.float32)[C]
    exp_pr, it = O.ordered_pagerank(R, C, W, oR, oC, np.ones(g.n, np.float32),
                                    np.ones(g.n, np.uint8),
                                    base=np.float32(1.0) - np.float32(0.85), damp=0.85, tol=1e-4)
    assert it == ig["iterations"]
    assert np.array_equal(pr_g.view(np.uint32), exp_pr.view(np.uint32))


def test_rejects_bad_input():
    case("ingest rejects")
    with pytest.raises(K.KitesError):
        K.DeviceGraph.from_edges(10, [1, 2], [3, 10], [1, 1])
    with pytest.raises(K.KitesError):
        K.DeviceGraph.from_edges(0, [], [], [])
    with pytest.raises(ValueError):
        K.DeviceGraph.from_edges(10, [1, 2], [3], [1, 1])
