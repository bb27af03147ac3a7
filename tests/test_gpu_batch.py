"""Multi-source runs (K-wide rows, up to 64 sources) against the oracle.

Column k of a batch must equal the single-source result from sources[k]:
Dijkstra for SSSP and BFS levels for BFS (both pinned in test_oracle.py),
bit for bit, for every work-item shape the autotuner may pick.
"""
import gzip
import os

import numpy as np
import pytest

from conftest import GOLDEN, TINY_GRAPHS, case, tiny_graph

pytestmark = pytest.mark.gpu

K = pytest.importorskip("kites_amd")
from oracle import liboracle as O  # noqa: E402


def pick_sources(oR, k, seed):
    outdeg = np.diff(oR.astype(np.int64))
    cand = np.flatnonzero(outdeg > 0)
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=min(k, len(cand)), replace=False).astype(np.uint32)


def expected(R, C, E, algo, sources):
    if algo == "sssp":
        return np.stack([O.dijkstra(R, C, E, int(s)) for s in sources], 1)
    return np.stack([O.bfs(R, C, int(s)) for s in sources], 1)


@pytest.fixture(scope="module")
def rmat14():
    h = K.HostGraph.rmat(14, 16, 41, wmin=1, wmax=255)
    return h, K.DeviceGraph(h)


@pytest.mark.parametrize("algo", ["sssp", "bfs"])
def test_64_sources_rmat14(rmat14, algo):
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = pick_sources(oR, 64, 1)
    case(f"batch {algo} rmat14 K=64")
    vals, info = dg.batch(algo, src, keep_stats=True)
    assert np.array_equal(vals, expected(R, C, E, algo, src))
    st = dg.batch_stats()
    assert len(st) == info["iterations"] and st[-1]["updated"] == 0
    assert all(s["gathered"] >= s["active_edges"] for s in st)


@pytest.mark.parametrize("algo", ["sssp", "bfs"])
def test_per_source_masks_equal_the_one_bit_frontier(rmat14, algo):
    """The per-source frontier masks (a 64-bit word per vertex of the sources
    whose value moved) give the same rows as the one-bit frontier, which
    equal Dijkstra / BFS per column (the variant is built for the default
    shape only; other shapes are refused)."""
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = pick_sources(oR, 64, 3)
    exp = expected(R, C, E, algo, src)
    case(f"batch {algo} rmat14 K=64 per-source masks")
    vals, _ = dg.batch(algo, src, per_source_masks=True)
    assert np.array_equal(vals, exp)
    with pytest.raises(K.KitesError):
        dg.batch(algo, src, per_source_masks=True, edges_per_item=128)


def test_partial_and_repeated_sources(rmat14):
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    hub = int(np.argmax(np.diff(oR)))
    src = np.array([hub, 5, hub, 16383, 0], np.uint32)
    case("batch sssp rmat14 K=5 with a repeated source")
    vals, _ = dg.batch("sssp", src)
    assert vals.shape == (dg.n, 5)
    assert np.array_equal(vals, expected(R, C, E, "sssp", src))
    assert np.array_equal(vals[:, 0], vals[:, 2])


@pytest.mark.parametrize("epi,block,unroll", [(128, 512, 16), (256, 256, 8), (512, 512, 8),
                                              (512, 256, 16)])
def test_shapes_agree(rmat14, epi, block, unroll):
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = pick_sources(oR, 64, 2)
    case(f"batch sssp rmat14 shape epi={epi} block={block} unroll={unroll}")
    vals, _ = dg.batch("sssp", src, edges_per_item=epi, block=block, unroll=unroll)
    assert np.array_equal(vals, expected(R, C, E, "sssp", src))


def test_without_graph_replay(rmat14):
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = pick_sources(oR, 64, 3)
    case("batch bfs rmat14 direct launches")
    a, _ = dg.batch("bfs", src, use_graph=False, chunk=1)
    b, _ = dg.batch("bfs", src, use_graph=True, chunk=5)
    assert np.array_equal(a, b) and np.array_equal(a, expected(R, C, E, "bfs", src))


def test_cached_graphs_replay_identical_runs():
    """Repeated K-wide and bit-parallel runs with other sources replay the
    captured chunks and stay exact."""
    h = K.HostGraph.rmat(12, 16, 43, wmin=1, wmax=255)
    dg = K.DeviceGraph(h)
    R, C, E, oR, oC, oE = h.arrays()
    case("batch cached hipGraph replays rmat12")
    _, i0 = dg.batch("sssp", pick_sources(oR, 64, 20))
    for seed in (21, 22):
        src = pick_sources(oR, 64, seed)
        vals, i1 = dg.batch("sssp", src)
        assert np.array_equal(vals, expected(R, C, E, "sssp", src))
        assert i1["graph_builds"] == i0["graph_builds"] >= 1
    _, b0 = dg.bfs256(pick_sources(oR, 200, 23))
    for seed in (24, 25):
        src = pick_sources(oR, 200, seed)
        lv, b1 = dg.bfs256(src)
        assert np.array_equal(lv, expected(R, C, E, "bfs", src))
        assert b1["graph_builds"] == b0["graph_builds"] > i0["graph_builds"]


def test_hub_cut_across_items():
    # vertex 0 receives an edge from every other vertex and sends one to
    # every other vertex: its in-list spans many work items (atomic flush).
    n = 20000
    others = np.arange(1, n, dtype=np.uint32)
    rng = np.random.default_rng(7)
    src = np.concatenate([others, np.zeros(n - 1, np.uint32), others[:-1]])
    dst = np.concatenate([np.zeros(n - 1, np.uint32), others, others[1:]])
    w = rng.integers(1, 1000, len(src)).astype(np.uint32)
    order = np.lexsort((src, dst))
    src, dst, w = src[order], dst[order], w[order]
    R = np.zeros(n + 1, np.int64)
    np.add.at(R, dst.astype(np.int64) + 1, 1)
    R = np.cumsum(R).astype(np.uint32)
    h = K.HostGraph.from_csr(R, src, w)
    R, C, E, *_ = h.arrays()
    dg = K.DeviceGraph(h)
    s = np.array([0, 1, 777, n - 1] + list(range(100, 160)), np.uint32)
    case("batch sssp hub graph n=20000 K=64")
    vals, _ = dg.batch("sssp", s, edges_per_item=128)
    assert np.array_equal(vals, expected(R, C, E, "sssp", s))


def test_wikivote_64_sources(tmp_path):
    p = tmp_path / "wiki.txt"
    p.write_bytes(gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read())
    h = K.HostGraph.load_edgelist(p)
    R, C, E, oR, oC, oE = h.arrays()
    dg = K.DeviceGraph(h)
    src = pick_sources(oR, 64, 4)
    case("batch sssp wiki-vote K=64")
    vals, _ = dg.batch("sssp", src)
    assert np.array_equal(vals, expected(R, C, E, "sssp", src))


def test_autotune_shapes_identical():
    h = K.HostGraph.rmat(12, 16, 43, wmin=1, wmax=255)
    R, C, E, oR, oC, oE = h.arrays()
    dg = K.DeviceGraph(h)
    src = pick_sources(oR, 64, 5)
    case("batch autotune rmat12")
    best, ms, table = dg.batch_autotune("sssp", src, repeats=1)
    assert len(table) == 12 and ms > 0
    assert best["edges_per_item"] in (128, 256, 512)
    vals, _ = dg.batch("sssp", src, **{k: best[k] for k in ("edges_per_item", "block", "unroll")})
    assert np.array_equal(vals, expected(R, C, E, "sssp", src))


@pytest.mark.parametrize("name", TINY_GRAPHS)
def test_tiny_graphs(name):
    """Both multi-source paths on tiny graphs, repeated sources included:
    each column equals Dijkstra / BFS from its source (an edgeless graph:
    the source alone at 0)."""
    h = tiny_graph(name)
    R, C, E, *_ = h.arrays()
    dg = K.DeviceGraph(h)
    n = dg.n
    src = np.array([0, n - 1, 0, n // 2, 1 % n, 40 % n], np.uint32)
    for algo in ("sssp", "bfs"):
        case(f"batch {algo} {name} K={len(src)}")
        vals, _ = dg.batch(algo, src)
        assert np.array_equal(vals, expected(R, C, E, algo, src)), algo
    for top_down in (1, 0):
        case(f"bfs256 {name} K={len(src)} top_down={top_down}")
        lev, info = dg.bfs256(src, top_down=top_down)
        assert np.array_equal(lev, expected(R, C, E, "bfs", src)), top_down
    case(f"batch and bfs256 {name} K=0")
    vals, info = dg.batch("sssp", np.zeros(0, np.uint32))
    assert vals.shape == (n, 0)
    lev, info = dg.bfs256(np.zeros(0, np.uint32))
    assert lev.shape == (n, 0)


def test_rejects_bad_arguments(rmat14):
    h, dg = rmat14
    with pytest.raises(K.KitesError):
        dg.batch("sssp", np.arange(65, dtype=np.uint32))
    with pytest.raises(K.KitesError):
        dg.batch("cc", np.arange(4, dtype=np.uint32))
    with pytest.raises(K.KitesError):
        dg.batch("sssp", np.array([dg.n], np.uint32))
    with pytest.raises(K.KitesError):
        dg.batch("sssp", np.arange(4, dtype=np.uint32), edges_per_item=100)
    with pytest.raises(K.KitesError):
        dg.batch("sssp", np.arange(4, dtype=np.uint32), block=128)


# ---- bit-parallel BFS (256 sources) -----------------------------------------

def test_bfs256_rmat14(rmat14):
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = pick_sources(oR, 256, 11)
    case("bfs256 rmat14 K=256")
    lev, info = dg.bfs256(src, keep_stats=True)
    assert lev.shape == (dg.n, 256)
    assert np.array_equal(lev, expected(R, C, E, "bfs", src))
    st = dg.bfs256_stats()
    assert len(st) == info["iterations"] and st[-1]["updated"] == 0
    reached = int((lev != K.INF).sum()) - len(src)
    assert sum(s["new_pairs"] for s in st) >= reached


@pytest.mark.parametrize("epi,block", [(64, 128), (512, 512), (4096, 256)])
def test_bfs256_shapes_and_partial(rmat14, epi, block):
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = np.concatenate([pick_sources(oR, 70, 12), [0, 0]]).astype(np.uint32)
    case(f"bfs256 rmat14 K=72 epi={epi} block={block}")
    lev, _ = dg.bfs256(src, edges_per_item=epi, block=block)
    assert np.array_equal(lev, expected(R, C, E, "bfs", src))


def test_bfs256_hub_and_wikivote(tmp_path):
    p = tmp_path / "wiki.txt"
    p.write_bytes(gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read())
    h = K.HostGraph.load_edgelist(p)
    R, C, E, oR, oC, oE = h.arrays()
    dg = K.DeviceGraph(h)
    src = pick_sources(oR, 256, 13)
    case("bfs256 wiki-vote K=256 epi=128")
    lev, _ = dg.bfs256(src, edges_per_item=128, use_graph=False, chunk=1)
    assert np.array_equal(lev, expected(R, C, E, "bfs", src))


def test_bfs256_deeper_than_a_byte():
    """Levels are kept as bytes while the run goes and stored straight as
    u32 from level 255 on: on a path of 700 vertices with two chords the
    levels reach 501, and the expanded u32 rows equal BFS exactly
    (unreached entries 0xFFFFFFFF).  Five vertices per level also pin the
    run's termination count (updated rows summed over every lane).  A second, shallow run on the same
    graph object reuses the byte rows and must not see the deep run's."""
    n = 700
    src_ids, dst_ids = list(range(n - 1)), list(range(1, n))
    src_ids += [10, 300]
    dst_ids += [200, 310]
    h = K.HostGraph.from_edges(n, np.array(src_ids, np.uint32), np.array(dst_ids, np.uint32),
                               np.ones(len(src_ids), np.uint32))
    R, C, E, oR, oC, oE = h.arrays()
    dg = K.DeviceGraph(h)
    src = np.array([0, 5, 299, 650, 699], np.uint32)
    case("bfs256 path of 700 vertices, 5 sources")
    lev, info = dg.bfs256(src)
    assert info["iterations"] > 255
    exp = expected(R, C, E, "bfs", src)
    assert int(exp[exp != K.INF].max()) == 501
    assert np.array_equal(lev, exp)
    src2 = np.array([690, 698], np.uint32)
    lev2, _ = dg.bfs256(src2)
    assert np.array_equal(lev2, expected(R, C, E, "bfs", src2))


@pytest.mark.parametrize("top_down", [0, 1, 2, 8])
def test_bfs256_top_down_levels(rmat14, top_down):
    """Levels whose frontier has few out-edges run top-down (push) unless
    top_down is 0; any direction rule gives BFS's levels exactly.  A
    top-down level claims each (vertex, source) pair once, so with every
    level top-down but the densest the pair counts are exact."""
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = pick_sources(oR, 256, 21)
    case(f"bfs256 rmat14 K=256 top_down={top_down}")
    lev, info = dg.bfs256(src, keep_stats=True, top_down=top_down)
    assert np.array_equal(lev, expected(R, C, E, "bfs", src))
    st = dg.bfs256_stats()
    pushed = [s["push"] for s in st]
    if top_down == 0:
        assert not any(pushed)
    else:
        assert any(pushed), pushed
        if top_down == 1:
            assert not all(pushed), pushed   # the dense levels pull
    reached = int((lev != K.INF).sum()) - len(src)
    assert sum(s["new_pairs"] for s in st) >= reached
    # a level's updated vertices and their out-edges steer the next level
    for s in st:
        if s["updated"] == 0:
            assert s["out_edges"] == 0


def test_bfs256_top_down_hub_and_path():
    """A hub whose out-list is longer than one wave may walk alone (4,096)
    makes its level pull; the long path before and after it pushes.  Levels
    equal BFS with the rule on and off, and the pull-only run's levels
    equal the top-down run's."""
    hub_leaves, path = 5000, 300
    n = 1 + hub_leaves + path
    src_ids = [0] * hub_leaves + [hub_leaves] + list(range(1 + hub_leaves, n - 1))
    dst_ids = list(range(1, 1 + hub_leaves)) + [1 + hub_leaves] + list(range(2 + hub_leaves, n))
    src_ids += [7, n - 1]
    dst_ids += [0, 3]
    h = K.HostGraph.from_edges(n, np.array(src_ids, np.uint32), np.array(dst_ids, np.uint32),
                               np.ones(len(src_ids), np.uint32))
    R, C, E, oR, oC, oE = h.arrays()
    dg = K.DeviceGraph(h)
    src = np.array([1 + hub_leaves, 7, 0, n - 50], np.uint32)
    exp = expected(R, C, E, "bfs", src)
    out = {}
    for top_down in (1, 0):
        case(f"bfs256 hub of {hub_leaves} and path of {path}, top_down={top_down}")
        lev, info = dg.bfs256(src, keep_stats=True, top_down=top_down)
        assert np.array_equal(lev, exp), top_down
        out[top_down] = [s["push"] for s in dg.bfs256_stats()]
    assert not out[1][0], "the hub's level pulls"
    assert any(out[1]) and not any(out[0])


def test_bfs256_autotune(rmat14):
    """Every (edges_per_item, block) shape the autotuner times gives the
    same levels (its device digest), and the tuned shape's levels are
    BFS's."""
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = pick_sources(oR, 200, 31)
    case("bfs256 autotune rmat14 K=200")
    best, ms, table = dg.bfs256_autotune(src, repeats=1)
    assert len(table) == 8 and ms > 0
    assert (best["edges_per_item"], best["block"]) in {(t[0], t[1]) for t in table}
    lev, _ = dg.bfs256(src, edges_per_item=best["edges_per_item"], block=best["block"])
    assert np.array_equal(lev, expected(R, C, E, "bfs", src))


def test_bfs256_matches_k_wide_bfs(rmat14):
    h, dg = rmat14
    R, C, E, oR, oC, oE = h.arrays()
    src = pick_sources(oR, 64, 14)
    case("bfs256 vs 64-wide bfs rmat14")
    a, _ = dg.bfs256(src)
    b, _ = dg.batch("bfs", src)
    assert np.array_equal(a, b)


# ---- uneven batch of independent graphs (one run, one source per graph) -----

def _sub(R, C, E, lo, hi):
    r = R[lo:hi + 1].astype(np.int64)
    return ((r - r[0]).astype(np.uint32), (C[r[0]:r[-1]] - lo).astype(np.uint32),
            E[r[0]:r[-1]].copy())


@pytest.mark.parametrize("sched", ["plain", "learned", "classic"])
def test_uneven_graph_batch(sched):
    g, voff, src = K.HostGraph.rmat_batch(60, 5, scale_lo=8, scale_hi=12, ef_lo=8, ef_hi=32)
    R, C, E, *_ = g.arrays()
    dg = K.DeviceGraph(g)
    if sched == "learned":
        from kites_amd import scheduler
        dg.set_model(scheduler.load_model())
    case(f"uneven batch 60 graphs sssp sched={sched}")
    dg.init_sources("sssp", src)
    info = dg.run("sssp", sched=sched)
    vals = dg.values()
    for i in range(60):
        lo, hi = int(voff[i]), int(voff[i + 1])
        r, c, e = _sub(R, C, E, lo, hi)
        assert np.array_equal(vals[lo:hi], O.dijkstra(r, c, e, int(src[i]) - lo)), i
    assert info["iterations"] >= 2


def test_graph_autotune_shapes_identical():
    g, voff, src = K.HostGraph.rmat_batch(30, 6, scale_lo=8, scale_hi=11, ef_lo=8, ef_hi=16)
    R, C, E, *_ = g.arrays()
    dg = K.DeviceGraph(g)
    case("single-source autotune on a 30-graph batch")
    best, ms, table = dg.autotune("sssp", src, sched="plain")
    assert len(table) == 24 and ms > 0 and best["grid"] > 0 and best["chunk"] in (4, 8, 16)
    dg.init_sources("sssp", src)
    dg.run("sssp", sched="plain", **best)
    vals = dg.values()
    for i in range(30):
        lo, hi = int(voff[i]), int(voff[i + 1])
        r, c, e = _sub(R, C, E, lo, hi)
        assert np.array_equal(vals[lo:hi], O.dijkstra(r, c, e, int(src[i]) - lo)), i
