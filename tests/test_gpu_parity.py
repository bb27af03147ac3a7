"""GPU parity of the engine against the CPU oracle, through the C ABI.

Every result is compared bit for bit with the oracle's Gauss-Seidel fixpoint
(the reference CPU path, kites/graph/sanity_check.cuh:21-47) run with the
same functor set on the same graph and initial state; the oracle itself is
pinned in tests/test_oracle.py.
"""
import gzip
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, TINY_GRAPHS, case, tiny_graph

pytestmark = pytest.mark.gpu

K = pytest.importorskip("kites_amd")
from oracle import liboracle as O  # noqa: E402

ORACLE_ALGO = {"sssp": O.SSSP_SAT, "bfs": O.BFS, "cc": O.CC, "sssp_wrap": O.SSSP_WRAP,
               "widest": O.WIDEST}

SCHEDULES = ([dict(sched="classic")] +
             [dict(sched="plain", gran=g) for g in (0, 1, 2, 3)] +
             [dict(sched="forced", forced_variant=v) for v in range(1, 9)])


def _expected(h, algo, src):
    R, C, E, oR, oC, oE = h.arrays()
    n = len(R) - 1
    if algo == "cc":
        V0, _ = O.cc_init(n)
    elif algo == "widest":
        V0, _ = O.widest_init(n, src, R, C)
    else:
        V0, _ = O.sssp_init(n, src, R, C)
        if algo == "sssp_wrap":
            V0[V0 == O.INF] = 1073741823
    V, _ = O.fixpoint(R, C, E, V0, ORACLE_ALGO[algo])
    return V


def _maxdeg_source(h):
    _, _, _, oR, _, _ = h.arrays()
    return int(np.argmax(np.diff(oR)))


@pytest.fixture(scope="module")
def c1():
    return K.HostGraph.rmat(13, 0, 1, wmin=1, wmax=64, n_limit=7115, exact_edges=103689)


@pytest.fixture(scope="module")
def c1_dev(c1):
    return K.DeviceGraph(c1)


@pytest.mark.parametrize("sch", SCHEDULES, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_c1_sssp_every_schedule(c1, c1_dev, sch):
    src = _maxdeg_source(c1)
    exp = _expected(c1, "sssp", src)
    case(f"C1 sssp src={src} {sch}")
    got, info = c1_dev.solve("sssp", src, **sch)
    assert info["iterations"] >= 2
    assert np.array_equal(got, exp), f"mismatches: {int((got != exp).sum())}"


@pytest.mark.parametrize("algo", ["bfs", "sssp_wrap", "widest"])
@pytest.mark.parametrize("sched", ["classic", "plain"])
def test_c1_other_functor_sets(c1, c1_dev, algo, sched):
    src = _maxdeg_source(c1)
    exp = _expected(c1, algo, src)
    case(f"C1 {algo} {sched}")
    got, info = c1_dev.solve(algo, src, sched=sched)
    assert np.array_equal(got, exp)
    assert info["delta"] == (sched != "classic")
    assert info["min_fast"] == (sched != "classic" and algo != "widest")


@pytest.mark.parametrize("sched", SCHEDULES[:5] + [dict(sched="forced", forced_variant=v) for v in (1, 4, 5, 8)],
                         ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_cc_symmetric(sched):
    h = K.HostGraph.rmat(12, 4, 3, wmin=1, wmax=9).symmetrize()
    d = K.DeviceGraph(h)
    exp = _expected(h, "cc", 0)
    R, C, *_ = h.arrays()
    assert np.array_equal(exp, O.cc_min_label(R, C))
    case(f"cc rmat12 {sched}")
    got, _ = d.solve("cc", 0, **sched)
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("sched", [dict(sched="classic"), dict(sched="plain", gran=1),
                                   dict(sched="plain", gran=3),
                                   dict(sched="forced", forced_variant=1),
                                   dict(sched="forced", forced_variant=4),
                                   dict(sched="forced", forced_variant=8)],
                         ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_grid_high_diameter(sched):
    h = K.HostGraph.grid(48, 40, 5, wmin=1, wmax=100)
    d = K.DeviceGraph(h)
    for algo in ("sssp", "bfs"):
        exp = _expected(h, algo, 0)
        case(f"grid48x40 {algo} {sched}")
        got, info = d.solve(algo, 0, **sched)
        assert np.array_equal(got, exp)
    # BFS on a grid: closed form |dr| + |dc| from the corner.
    rr, cc = np.divmod(np.arange(48 * 40), 40)
    assert np.array_equal(got[:48 * 40], (rr + cc).astype(np.uint32))


def test_hipgraph_equals_direct_launches(c1, c1_dev):
    src = _maxdeg_source(c1)
    exp = _expected(c1, "sssp", src)
    case("C1 sssp without hipGraph")
    a, ia = c1_dev.solve("sssp", src, sched="plain", use_graph=False, chunk=3)
    b, ib = c1_dev.solve("sssp", src, sched="plain", use_graph=True, chunk=7)
    assert np.array_equal(a, exp) and np.array_equal(b, exp)


def test_cached_graphs_replay_identical_runs(c1):
    """Repeated runs replay the captured iteration chunks (no new capture)
    and give the oracle's result every time; another schedule or source
    functor captures its own chunks."""
    d = K.DeviceGraph(c1)
    src = _maxdeg_source(c1)
    exp = {a: _expected(c1, a, src) for a in ("sssp", "bfs")}
    case("C1 cached hipGraph replays")
    v, i0 = d.solve("sssp", src, sched="plain")
    assert np.array_equal(v, exp["sssp"]) and i0["graph_builds"] >= 1
    for other in (src, 0, 17):
        # the source is an argument of the seed kernel, not of the step kernel
        v, i1 = d.solve("sssp", other, sched="plain")
        assert i1["graph_builds"] == i0["graph_builds"]
        assert np.array_equal(v, _expected(c1, "sssp", other) if other != src else exp["sssp"])
    v, i2 = d.solve("sssp", src, sched="forced", forced_variant="push_lb")
    assert np.array_equal(v, exp["sssp"]) and i2["graph_builds"] > i0["graph_builds"]
    v, i3 = d.solve("bfs", src, sched="plain")
    assert np.array_equal(v, exp["bfs"]) and i3["graph_builds"] > i2["graph_builds"]
    for _ in range(2):
        v, i4 = d.solve("sssp", src, sched="plain")
        assert np.array_equal(v, exp["sssp"]) and i4["graph_builds"] == i3["graph_builds"]
        v, i5 = d.solve("bfs", src, sched="plain")
        assert np.array_equal(v, exp["bfs"]) and i5["graph_builds"] == i3["graph_builds"]


def test_edge_cases():
    # no edges at all
    R = np.zeros(129, np.uint32)
    h = K.HostGraph.from_csr(R, np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    d = K.DeviceGraph(h)
    case("edgeless graph")
    got, info = d.solve("sssp", 5)
    exp = np.full(128, K.INF, np.uint32)
    exp[5] = 0
    assert np.array_equal(got, exp)
    assert info["iterations"] == 1
    # star with a hub (one source to everyone) and a path behind it
    n = 256
    src_list = [0] * 200 + list(range(200, 255))
    dst_list = list(range(1, 201)) + list(range(201, 256))
    order = np.lexsort((src_list, dst_list))
    s = np.array(src_list)[order]
    t = np.array(dst_list)[order]
    w = (np.arange(len(s)) % 7 + 1).astype(np.uint32)
    R = np.zeros(n + 1, np.uint32)
    np.add.at(R, t + 1, 1)
    R = np.cumsum(R).astype(np.uint32)
    h = K.HostGraph.from_csr(R, s.astype(np.uint32), w)
    d = K.DeviceGraph(h)
    for sch in SCHEDULES:
        exp = _expected(h, "sssp", 0)
        case(f"star+path {sch}")
        got, _ = d.solve("sssp", 0, **sch)
        assert np.array_equal(got, exp), sch
    # isolated source: nothing but the source changes
    case("isolated source")
    got, info = d.solve("sssp", 255)
    exp = np.full(n, K.INF, np.uint32)
    exp[255] = 0
    assert np.array_equal(got, exp)


TINY_SCHEDULES = SCHEDULES + [dict(sched="learned"), dict(sched="plain", persist=False),
                              dict(sched="learned", persist=False),
                              dict(sched="plain", split_pull=True),
                              dict(sched="forced", forced_variant="pull_lb", split_pull=True)]
TINY_DELTA = [dict(sched="plain", delta=4), dict(sched="learned", delta=4),
              dict(sched="forced", forced_variant="push_lb", delta=4),
              dict(sched="plain", delta=4, persist=False)]


@pytest.mark.parametrize("name", TINY_GRAPHS)
def test_tiny_graphs_every_schedule(name):
    """Every schedule and variant (and near-far for SSSP) on the tiny graphs,
    for every functor set, from the first and the last vertex and from one
    on a zero-weight cycle: the oracle's fixpoint bit for bit."""
    from kites_amd import scheduler
    h = tiny_graph(name)
    d = K.DeviceGraph(h)
    d.set_model(scheduler.load_model())
    n = d.n
    for algo in ("sssp", "bfs", "widest", "cc"):
        for src in sorted({0, n - 1, 40 % n}):
            exp = _expected(h, algo, src)
            for sch in TINY_SCHEDULES + (TINY_DELTA if algo == "sssp" else []):
                case(f"tiny {name} {algo} from {src} {sch}")
                got, _ = d.solve(algo, src, **sch)
                assert np.array_equal(got, exp), (algo, src, sch)


@pytest.mark.parametrize("name", ("C1",) + TINY_GRAPHS)
def test_convenience_entry_points(name, c1):
    """kites_sssp_u32, kites_bfs_u32 and kites_cc_u32 called through their
    own symbols (default options as NULL and as a filled struct), a
    multi-source start (kites_graph_init_sources: the elementwise minimum
    of the single-source fixpoints) and kites_pagerank_f32 (bit for bit the
    oracle's ordered PageRank), on C1 and on the tiny graphs."""
    import ctypes
    from kites_amd import _lib
    h = c1 if name == "C1" else tiny_graph(name)
    R, C, E, oR, oC, oE = h.arrays()
    d = K.DeviceGraph(h)
    n = d.n
    L = _lib.lib()
    info = _lib.RunInfo()
    out = np.empty(n, np.uint32)
    p = out.ctypes.data_as(_lib.u32p)
    src = _maxdeg_source(h)
    filled = _lib.RunOpts()
    L.kites_run_opts_default(ctypes.byref(filled))
    for opts in (None, ctypes.byref(filled)):
        tag = "NULL" if opts is None else "default struct"
        case(f"{name} kites_sssp_u32 / kites_bfs_u32 / kites_cc_u32 from {src}, options {tag}")
        _lib.check(L.kites_sssp_u32(d._g, src, opts, ctypes.byref(info), p))
        assert np.array_equal(out, _expected(h, "sssp", src))
        _lib.check(L.kites_bfs_u32(d._g, src, opts, ctypes.byref(info), p))
        assert np.array_equal(out, _expected(h, "bfs", src))
        _lib.check(L.kites_cc_u32(d._g, opts, ctypes.byref(info), p))
        assert np.array_equal(out, _expected(h, "cc", 0))
    with pytest.raises(K.KitesError):
        _lib.check(L.kites_sssp_u32(d._g, n, None, ctypes.byref(info), p))
    sources = np.unique(np.array([0, n - 1, n // 2, src], np.uint32))
    for algo in ("sssp", "bfs"):
        case(f"{name} {algo} from {len(sources)} sources at once")
        d.init_sources(algo, sources)
        d.run(algo, sched="plain")
        exp = np.minimum.reduce([_expected(h, algo, int(s)) for s in sources])
        assert np.array_equal(d.values(), exp), algo
    case(f"{name} kites_pagerank_f32")
    outdeg = np.diff(oR.astype(np.int64)).astype(np.float32)
    W = np.float32(1.0) / outdeg[C]
    got, pinfo = d.pagerank(damping=0.85, tol=1e-4)
    exp, it = O.ordered_pagerank(R, C, W, oR, oC, np.ones(n, np.float32), np.ones(n, np.uint8),
                                 base=np.float32(1.0) - np.float32(0.85), damp=0.85, tol=1e-4)
    assert pinfo["iterations"] == it
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    d.close()


def test_set_state_generic_initial_frontier(c1, c1_dev):
    """Arbitrary initial values and frontier (the reference callback's
    protocol) give the oracle's fixpoint when the frontier covers every
    vertex whose in-edges disagree with its value."""
    R, C, E, *_ = c1.arrays()
    n = len(R) - 1
    rng = np.random.default_rng(7)
    V0 = np.full(n, K.INF, np.uint32)
    seeds = rng.choice(n, 5, replace=False)
    V0[seeds] = rng.integers(0, 50, 5).astype(np.uint32)
    F0 = np.zeros(n, np.uint8)
    dst_of_edge = np.repeat(np.arange(n), np.diff(R).astype(np.int64))
    F0[dst_of_edge[np.isin(C, seeds)]] = 1
    exp, _ = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
    for sch in ("classic", "plain"):
        c1_dev.set_state(V0, F0)
        case(f"C1 set_state {sch}")
        c1_dev.run("sssp", sched=sch)
        assert np.array_equal(c1_dev.values(), exp)


@pytest.mark.parametrize("scale", [16, 18])
def test_rmat_sssp_bfs_plain(scale):
    h = K.HostGraph.rmat(scale, 16, 11, wmin=1, wmax=255)
    d = K.DeviceGraph(h)
    src = _maxdeg_source(h)
    R, C, E, *_ = h.arrays()
    for algo in ("sssp", "bfs"):
        exp = O.dijkstra(R, C, E, src) if algo == "sssp" else O.bfs(R, C, src)
        for g in (0, 1, 2, 3):
            case(f"rmat{scale} {algo} plain gran={g}")
            got, info = d.solve(algo, src, sched="plain", gran=g)
            assert np.array_equal(got, exp), f"{algo} gran {g}: {int((got != exp).sum())} diffs"


def test_wikivote_dropin_example(tmp_path):
    """The reference example, compiled against the drop-in C++ API, on the
    reference's own data file; result equals the committed golden vector
    (and the example's sanity_check passes)."""
    exe = os.path.join(ROOT, "examples", "_build", "sssp_basic_single_device")
    assert os.path.exists(exe), "examples not built"
    txt = tmp_path / "wiki.txt"
    txt.write_bytes(gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read())
    golden = np.load(os.path.join(GOLDEN, "wikivote_sssp_src30.npz"))["V"]
    for vpg in (1, 4, 32):
        out = tmp_path / f"v{vpg}.u32"
        case(f"wikivote example vpg={vpg}")
        r = subprocess.run([exe, str(txt), "30", str(vpg), str(out)], capture_output=True,
                           text=True, timeout=120)
        print(r.stdout[-2000:], r.stderr[-2000:])
        assert r.returncode == 0
        assert "SANITY CHECK: PASSED" in r.stdout
        got = np.fromfile(out, np.uint32)
        assert np.array_equal(got, golden)


def test_vertex_per_group_autotune(tmp_path):
    """The vertex-per-group knob of the reference algorithm: every vpg the
    tuner tries gives the golden vector (Wiki-Vote, source 30, the example's
    functors), and it returns the fastest."""
    txt = tmp_path / "wiki.txt"
    txt.write_bytes(gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read())
    golden = np.load(os.path.join(GOLDEN, "wikivote_sssp_src30.npz"))["V"]
    case("vertex-per-group autotune on wiki-vote")
    best, ms, table = K.autotune_vpg(txt, source=30, vpgs=(1, 2, 4, 8, 16, 32, 64), repeats=2)
    print("vpg table (ms):", table, "best", best, flush=True)
    assert set(table) == {1, 2, 4, 8, 16, 32, 64} and table[best] == ms == min(table.values())
    h = K.HostGraph.load_edgelist(txt, vpg=best, source=30)
    V0, F0 = h.load_arrays()[:2]
    d = K.DeviceGraph(h)
    d.set_state(V0, F0)
    d.run("sssp_wrap", sched="classic")
    assert np.array_equal(d.values(), golden)


def test_host_graph_process_overload(tmp_path):
    """process<launch::sync>(graph_csr&, gpu&, F x 4), the host-graph
    overload (kites/device/compute.cuh:397-411): upload, run and copy back
    into the host graph's V, which equals the committed golden vector."""
    exe = os.path.join(ROOT, "examples", "_build", "sssp_host_graph")
    assert os.path.exists(exe), "examples not built"
    txt = tmp_path / "wiki.txt"
    txt.write_bytes(gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read())
    golden = np.load(os.path.join(GOLDEN, "wikivote_sssp_src30.npz"))["V"]
    out = tmp_path / "v.u32"
    case("wikivote host-graph process overload")
    r = subprocess.run([exe, str(txt), "30", str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "Performed" in r.stdout and "Source value before 0, after 0" in r.stdout
    got = np.fromfile(out, np.uint32)
    assert np.array_equal(got, golden)
    assert int((golden != 1073741823).sum()) > 1000


def test_prepared_graph_dropin(tmp_path):
    """kites::io::save_to_file + graph_csr(file) of the drop-in API: the dump
    equals the C ABI's text dump of the same edge list byte for byte (both
    pinned by the oracle's restatement of the reference dump), and SSSP on
    the reloaded graph equals the golden vector."""
    exe = os.path.join(ROOT, "examples", "_build", "prepared_graph")
    assert os.path.exists(exe), "examples not built"
    txt = tmp_path / "wiki.txt"
    txt.write_bytes(gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read())
    golden = np.load(os.path.join(GOLDEN, "wikivote_sssp_src30.npz"))["V"]
    dump, out, capi = tmp_path / "dump.txt", tmp_path / "v.u32", tmp_path / "capi.txt"
    case("prepared graph drop-in")
    r = subprocess.run([exe, str(txt), str(dump), str(out), "30"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "SANITY CHECK: PASSED" in r.stdout
    K.HostGraph.load_edgelist(txt, source=30).save_text(capi)
    assert dump.read_bytes() == capi.read_bytes()
    assert np.array_equal(np.fromfile(out, np.uint32), golden)


def test_ordered_pagerank_dropin(tmp_path):
    """A non-idempotent, order-sensitive functor set (float PageRank) through
    the drop-in API with options::ordered: bit for bit the oracle's ordered
    Jacobi iterations with the same combine order."""
    from oracle import graph_raw_oracle as gro
    exe = os.path.join(ROOT, "examples", "_build", "pagerank_ordered")
    assert os.path.exists(exe), "examples not built"
    raw = gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read()
    txt = tmp_path / "wiki.txt"
    txt.write_bytes(raw)
    out = tmp_path / "pr.f32"
    case("ordered pagerank drop-in")
    r = subprocess.run([exe, str(txt), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    got = np.fromfile(out, np.float32)
    g = gro.build(raw.decode(), lambda argv, s, d, V: (0, True), 1, True, True, True)
    R, C, oR, oC = g["R"], g["C"], g["dir_R"], g["dir_C"]
    n = len(R) - 1
    outdeg = np.diff(oR.astype(np.int64)).astype(np.float32)
    W = np.float32(1.0) / outdeg[C]
    seen = np.zeros(n, bool)
    seen[C] = True
    seen[np.repeat(np.arange(n), np.diff(R.astype(np.int64)))] = True
    V0 = np.where(seen, np.float32(1.0), np.float32(0.0)).astype(np.float32)
    exp, it = O.ordered_pagerank(R, C, W, oR, oC, V0, g["frontier_as_reference"])
    assert it > 3 and len(got) == n
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), \
        f"differ at {int((got != exp).sum())} vertices"


@pytest.mark.parametrize("scale,damping,tol", [(12, 0.85, 1e-4), (16, 0.85, 1e-4), (15, 0.5, 0.0)])
def test_pagerank_capi(scale, damping, tol):
    """kites_pagerank_f32 (the ordered mode behind the C ABI) on R-MAT graphs:
    values and iteration count bit for bit the oracle's ordered Jacobi
    PageRank, every vertex starting at 1 and active."""
    case(f"pagerank capi scale {scale} damping {damping} tol {tol}")
    h = K.HostGraph.rmat(scale, 16, 40 + scale, wmin=1, wmax=9)
    R, C, _, oR, oC, _ = h.arrays()
    n = len(R) - 1
    outdeg = np.diff(oR.astype(np.int64)).astype(np.float32)
    W = np.float32(1.0) / outdeg[C]
    d = K.DeviceGraph(h)
    got, info = d.pagerank(damping=damping, tol=tol)
    exp, it = O.ordered_pagerank(R, C, W, oR, oC, np.ones(n, np.float32), np.ones(n, np.uint8),
                                 base=np.float32(1.0) - np.float32(damping), damp=damping,
                                 tol=tol)
    assert it > 3 and info["iterations"] == it
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), \
        f"differ at {int((got != exp).sum())} vertices"
    # the graph's own state is untouched by the float view
    again, info2 = d.pagerank(damping=damping, tol=tol)
    assert info2["iterations"] == it and np.array_equal(again.view(np.uint32), got.view(np.uint32))
    with pytest.raises(RuntimeError, match="damping"):
        d.pagerank(damping=1.5)


# ---- near-far push (delta-bucketed expansion) --------------------------------

@pytest.mark.parametrize("delta", [1, 64, 300])
@pytest.mark.parametrize("sch", [dict(sched="plain"), dict(sched="forced", forced_variant="push_t"),
                                 dict(sched="forced", forced_variant="push_w"),
                                 dict(sched="forced", forced_variant="push_b"),
                                 dict(sched="forced", forced_variant="push_lb"),
                                 dict(sched="plain", gran=0)],
                         ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_near_far_sssp(delta, sch):
    h = K.HostGraph.rmat(15, 16, 77, wmin=1, wmax=255)
    R, C, E, oR, oC, oE = h.arrays()
    src = _maxdeg_source(h)
    case(f"near-far sssp rmat15 delta={delta} {sch}")
    dg = K.DeviceGraph(h)
    vals, info = dg.solve("sssp", src, delta=delta, **sch)
    assert np.array_equal(vals, O.dijkstra(R, C, E, src))


@pytest.mark.parametrize("delta", [16, 255])
@pytest.mark.parametrize("sch", [dict(sched="learned"), dict(sched="forced", forced_variant="pull_lb"),
                                 dict(sched="forced", forced_variant="pull_t"),
                                 dict(sched="plain", alpha=1000.0)],
                         ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_near_far_with_pulls_and_learned(delta, sch):
    """Pull iterations queue every vertex they update, whatever its value;
    the far pile keeps its entries across them and still drains."""
    h = K.HostGraph.rmat(16, 16, 79, wmin=1, wmax=255)
    R, C, E, oR, oC, oE = h.arrays()
    src = _maxdeg_source(h)
    dg = K.DeviceGraph(h)
    if sch["sched"] == "learned":
        from kites_amd import scheduler
        dg.set_model(scheduler.load_model())
    case(f"near-far sssp rmat16 delta={delta} {sch}")
    vals, info = dg.solve("sssp", src, delta=delta, keep_stats=True, **sch)
    assert np.array_equal(vals, O.dijkstra(R, C, E, src))
    st = dg.stats()
    assert len(st) == info["iterations"]


def test_near_far_split_and_work():
    """A small delta makes the queue run dry with entries on the far pile:
    split iterations appear, the run still equals Dijkstra, and it walks no
    more edges than the run without near-far."""
    h = K.HostGraph.rmat(17, 16, 80, wmin=1, wmax=255)
    R, C, E, oR, oC, oE = h.arrays()
    src = _maxdeg_source(h)
    dg = K.DeviceGraph(h)
    exp = O.dijkstra(R, C, E, src)
    work = {}
    for delta in (0, 32):
        case(f"near-far work rmat17 delta={delta}")
        vals, info = dg.solve("sssp", src, delta=delta, sched="forced", forced_variant="push_lb",
                              keep_stats=True)
        assert np.array_equal(vals, exp)
        st = dg.stats()
        work[delta] = sum(s["work_edges"] for s in st)
        splits = sum(1 for s in st if K.VARIANTS[s["variant"]] == "split")
        print(f"delta {delta}: {info['iterations']} iterations, {splits} splits, "
              f"{work[delta]} edges walked", flush=True)
        if delta:
            assert splits >= 1
    assert work[32] <= work[0]


def test_near_far_grid_high_diameter():
    h = K.HostGraph.grid(96, 80, 5, wmin=1, wmax=100)
    R, C, E, oR, oC, oE = h.arrays()
    src = 0
    dg = K.DeviceGraph(h)
    case("near-far sssp grid 96x80 delta=50")
    vals, _ = dg.solve("sssp", src, delta=50, sched="plain")
    assert np.array_equal(vals, O.dijkstra(R, C, E, src))


@pytest.mark.parametrize("algo", ["bfs", "sssp_wrap"])
def test_near_far_other_sets(algo):
    h = K.HostGraph.rmat(14, 8, 78, wmin=1, wmax=64)
    case(f"near-far {algo} rmat14 delta=2")
    dg = K.DeviceGraph(h)
    src = _maxdeg_source(h)
    vals, _ = dg.solve(algo, src, delta=2, sched="plain")
    assert np.array_equal(vals, _expected(h, algo, src))


# ---- source-split pull (one part of the sources per XCD) --------------------

SPLIT = [dict(sched="forced", forced_variant="pull_lb", split_pull=True),
         dict(sched="plain", split_pull=True)]
_sid = lambda d: "-".join(f"{k}{v}" for k, v in d.items())  # noqa: E731


@pytest.mark.parametrize("algo", ["sssp", "bfs", "sssp_wrap", "widest"])
@pytest.mark.parametrize("sch", SPLIT, ids=_sid)
def test_split_pull_c1(c1, c1_dev, algo, sch):
    """The edge-balanced pull over the source-split in-CSR gives the oracle's
    fixpoint for every functor set (widest runs the generic CAS relax)."""
    src = _maxdeg_source(c1)
    exp = _expected(c1, algo, src)
    case(f"C1 split pull {algo} {sch}")
    got, info = c1_dev.solve(algo, src, **sch)
    assert np.array_equal(got, exp), f"mismatches: {int((got != exp).sum())}"


@pytest.mark.parametrize("scale", [14, 17])
@pytest.mark.parametrize("sch", SPLIT, ids=_sid)
def test_split_pull_rmat(scale, sch):
    h = K.HostGraph.rmat(scale, 16, 90 + scale, wmin=1, wmax=255)
    d = K.DeviceGraph(h)
    src = _maxdeg_source(h)
    for algo in ("sssp", "bfs"):
        exp = _expected(h, algo, src)
        case(f"rmat{scale} split pull {algo} {sch}")
        got, info = d.solve(algo, src, **sch)
        assert np.array_equal(got, exp), f"mismatches: {int((got != exp).sum())}"
        # the same graph without the split (the split stays built beside it)
        plain, _ = d.solve(algo, src, sched="forced", forced_variant="pull_lb")
        assert np.array_equal(plain, exp)


def test_split_pull_cc_and_grid():
    h = K.HostGraph.rmat(12, 4, 3, wmin=1, wmax=9).symmetrize()
    d = K.DeviceGraph(h)
    exp = _expected(h, "cc", 0)
    for sch in SPLIT:
        case(f"cc rmat12 split {sch}")
        got, _ = d.solve("cc", 0, **sch)
        assert np.array_equal(got, exp)
    g = K.HostGraph.grid(48, 40, 5, wmin=1, wmax=100)
    dg = K.DeviceGraph(g)
    for sch in SPLIT:
        exp = _expected(g, "sssp", 0)
        case(f"grid48x40 split {sch}")
        got, _ = dg.solve("sssp", 0, **sch)
        assert np.array_equal(got, exp)


def test_split_pull_small_graphs():
    """Fewer edges than parts, an edgeless graph, and a hub whose in-edges
    all come from one part."""
    case("split pull edgeless")
    h = K.HostGraph.from_csr(np.zeros(129, np.uint32), np.zeros(0, np.uint32),
                             np.zeros(0, np.uint32))
    d = K.DeviceGraph(h)
    for sch in SPLIT:
        got, info = d.solve("sssp", 5, **sch)
        exp = np.full(128, K.INF, np.uint32)
        exp[5] = 0
        assert np.array_equal(got, exp)
    case("split pull three edges")
    # 0 -> 1 (w 4), 1 -> 2 (w 1), 0 -> 2 (w 9): in-lists by destination
    R = np.zeros(129, np.uint32)
    R[2:] = 1
    R[3:] = 3
    h = K.HostGraph.from_csr(R, np.array([0, 0, 1], np.uint32), np.array([4, 9, 1], np.uint32))
    d = K.DeviceGraph(h)
    for sch in SPLIT:
        got, _ = d.solve("sssp", 0, **sch)
        assert got[:3].tolist() == [0, 4, 5] and (got[3:] == K.INF).all()
    case("split pull hub fed by one part")
    n = 1024
    srcs = np.arange(32, dtype=np.uint32)          # one 32-vertex line: one part
    dst = np.full(32, 1000, np.uint32)
    chain_s = np.arange(1000, 1023, dtype=np.uint32)
    chain_d = chain_s + 1
    s_all = np.concatenate([srcs, chain_s, np.zeros(31, np.uint32)])
    d_all = np.concatenate([dst, chain_d, np.arange(1, 32, dtype=np.uint32)])
    order = np.lexsort((s_all, d_all))
    s_all, d_all = s_all[order], d_all[order]
    w = ((s_all * 7 + d_all) % 13 + 1).astype(np.uint32)
    R = np.zeros(n + 1, np.uint32)
    np.add.at(R, d_all.astype(np.int64) + 1, 1)
    R = np.cumsum(R).astype(np.uint32)
    h = K.HostGraph.from_csr(R, s_all, w)
    d = K.DeviceGraph(h)
    exp = _expected(h, "sssp", 0)
    for sch in SPLIT:
        got, _ = d.solve("sssp", 0, **sch)
        assert np.array_equal(got, exp)
