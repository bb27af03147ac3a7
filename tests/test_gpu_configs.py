"""GPU parity at the workloads' own sizes, every vertex checked.

* R-MAT scale 22, edge factor 16 (4,194,304 vertices, 65.2M edges, seed 22,
  weights U[1,255]): plain SSSP against textbook Dijkstra and plain BFS
  against textbook BFS from the max-out-degree vertex.
* The learned schedule on R-MAT-22, Erdos-Renyi (n = 4,194,304, 16n draws,
  seed 201) and the 2048 x 2048 4-neighbour grid (weights U[1,100], seed
  201): SSSP, BFS and min-label CC (R-MAT and ER symmetrized) bit-equal to
  the plain schedule and to the oracle (Dijkstra, BFS, union-find).
* 64-source SSSP and 256-source bit-parallel BFS on R-MAT-22: every column
  through the shortest-path certificate (value 0 at the source; no edge
  relaxes any entry; every other finite entry has a tight in-edge; positive
  weights make the certificate's solution unique, so passing it means
  equal to Dijkstra), and 4 columns against Dijkstra / BFS directly.
* 1,000 uneven R-MAT graphs (scale U{10..16}, edge factor U{8..32}, seed
  1000) in one run, learned (at the autotuned shape) and plain, against
  per-graph Dijkstra.

The reference checks its whole result on every run the same way
(kites/graph/sanity_check.cuh:21-57, called at
examples/sssp_basic_single_device.cu:69): only final values are compared.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import case

pytestmark = pytest.mark.gpu

K = pytest.importorskip("kites_amd")
from kites_amd import scheduler as S  # noqa: E402
from oracle import liboracle as O  # noqa: E402

_POOL = ThreadPoolExecutor(max_workers=4)   # oracle calls release the GIL


def _maxdeg_source(oR):
    return int(np.argmax(np.diff(oR.astype(np.int64))))


class _Workload:
    """A host graph, its device copy and oracle results computed in the
    background (so the oracle's CPU time overlaps the GPU runs)."""

    def __init__(self, name, h, symmetrize_cc=True):
        self.name, self.h = name, h
        self.R, self.C, self.E, self.oR, self.oC, self.oE = h.arrays()
        self.src = _maxdeg_source(self.oR)
        self.d = K.DeviceGraph(h)
        self.d.set_model(S.load_model())
        self.fut = {
            "sssp": _POOL.submit(O.dijkstra, self.R, self.C, self.E, self.src),
            "bfs": _POOL.submit(O.bfs, self.R, self.C, self.src),
        }
        self.symmetrize_cc = symmetrize_cc
        self._cc = None

    def expected(self, algo):
        return self.fut[algo].result()

    def cc(self):
        """(device graph, union-find labels) of the graph CC runs on."""
        if self._cc is None:
            if self.symmetrize_cc:
                hs = self.h.symmetrize()
            else:
                hs = self.h
            R, C = hs.arrays()[:2]
            fut = _POOL.submit(O.cc_min_label, R, C)
            ds = self.d if hs is self.h else K.DeviceGraph(hs)
            ds.set_model(S.load_model())
            self._cc = (ds, fut)
        return self._cc

    def close(self):
        self.d.close()
        if self._cc is not None and self._cc[0] is not self.d:
            self._cc[0].close()


@pytest.fixture(scope="module")
def rmat22():
    w = _Workload("rmat22", K.HostGraph.rmat(22, 16, 22, wmin=1, wmax=255))
    yield w
    w.close()


@pytest.fixture(scope="module")
def er22():
    w = _Workload("er22", K.HostGraph.erdos_renyi(1 << 22, 16 << 22, 201))
    yield w
    w.close()


@pytest.fixture(scope="module")
def grid2048():
    h = K.HostGraph.grid(2048, 2048, 201, wmin=1, wmax=100)
    R, C = h.arrays()[:2]
    # every edge has its reverse: CC runs on the grid itself
    dst = np.repeat(np.arange(len(R) - 1, dtype=np.uint64), np.diff(R.astype(np.int64)))
    fwd = np.sort(C.astype(np.uint64) << np.uint64(32) | dst)
    rev = np.sort(dst << np.uint64(32) | C.astype(np.uint64))
    assert np.array_equal(fwd, rev), "grid is not structurally symmetric"
    w = _Workload("grid2048", h, symmetrize_cc=False)
    yield w
    w.close()


def test_rmat22_sizes(rmat22):
    assert rmat22.d.n == 4194304 and len(rmat22.R) - 1 == 4194304
    assert 60_000_000 < rmat22.d.m <= 67_108_864
    assert rmat22.E.min() >= 1 and rmat22.E.max() <= 255


@pytest.mark.parametrize("algo", ["sssp", "bfs"])
def test_rmat22_plain_single_source(rmat22, algo):
    case(f"rmat22 plain {algo} src={rmat22.src}")
    got, info = rmat22.d.solve(algo, rmat22.src, sched="plain")
    exp = rmat22.expected(algo)
    assert np.array_equal(got, exp), f"{int((got != exp).sum())} vertices differ"
    assert info["delta"] and info["min_fast"]
    # about half of R-MAT-22's ids have no edges at all
    assert (got != K.INF).sum() > 0.4 * rmat22.d.n


def _learned_vs_plain(w, algo):
    if algo == "cc":
        d, fut = w.cc()
        src = 0
    else:
        d, fut, src = w.d, w.fut[algo], w.src
    case(f"{w.name} {algo} plain/learned src={src}")
    plain, ip = d.solve(algo, src, sched="plain")
    learned, il = d.solve(algo, src, sched="learned")
    assert np.array_equal(learned, plain), f"{int((learned != plain).sum())} vertices differ"
    exp = fut.result()
    assert np.array_equal(plain, exp), f"{int((plain != exp).sum())} vertices differ from the oracle"
    print(f"{w.name} {algo}: plain {ip['ms']:.2f} ms / {ip['iterations']} it, "
          f"learned {il['ms']:.2f} ms / {il['iterations']} it", flush=True)


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_learned_rmat22(rmat22, algo):
    _learned_vs_plain(rmat22, algo)


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_learned_er22(er22, algo):
    _learned_vs_plain(er22, algo)


@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_learned_grid2048(grid2048, algo):
    _learned_vs_plain(grid2048, algo)


def test_grid2048_closed_form_bfs(grid2048):
    # BFS levels on the 4-neighbour grid are Manhattan distances
    w = grid2048
    lev = w.expected("bfs")
    r0, c0 = divmod(w.src, 2048)
    rr, cc = np.divmod(np.arange(2048 * 2048), 2048)
    assert np.array_equal(lev[:2048 * 2048], (np.abs(rr - r0) + np.abs(cc - c0)).astype(np.uint32))


def _sources(oR, k, seed):
    cand = np.flatnonzero(np.diff(oR.astype(np.int64)) > 0)
    return np.random.default_rng(seed).choice(cand, k, replace=False).astype(np.uint32)


def test_rmat22_64_sources(rmat22):
    w = rmat22
    src = _sources(w.oR, 64, 64)
    direct = [_POOL.submit(O.dijkstra, w.R, w.C, w.E, int(src[k])) for k in (0, 21, 42, 63)]
    case("rmat22 64-source sssp")
    rows, info = w.d.batch("sssp", src)
    assert rows.shape == (w.d.n, 64)
    tot, bad = O.check_cols(w.R, w.C, w.E, rows, src)
    assert tot == 0, f"certificate: {bad} (bad sources, relaxable edges, untight entries)"
    for k, f in zip((0, 21, 42, 63), direct):
        assert np.array_equal(rows[:, k], f.result()), k
    # the single-source run of the first column equals it too
    assert np.array_equal(rows[:, 0], w.d.solve("sssp", int(src[0]), sched="plain")[0])


def test_rmat22_256_source_bfs(rmat22):
    w = rmat22
    src = _sources(w.oR, 256, 256)
    direct = [_POOL.submit(O.bfs, w.R, w.C, int(src[k])) for k in (0, 85, 170, 255)]
    case("rmat22 256-source bit-parallel bfs")
    lev, info = w.d.bfs256(src)
    assert lev.shape == (w.d.n, 256)
    tot, bad = O.check_cols(w.R, w.C, None, lev, src)
    assert tot == 0, f"certificate: {bad} (bad sources, relaxable edges, untight entries)"
    for k, f in zip((0, 85, 170, 255), direct):
        assert np.array_equal(lev[:, k], f.result()), k
    del lev


@pytest.fixture(scope="module")
def graphs1000():
    g, voff, src = K.HostGraph.rmat_batch(1000, 1000, scale_lo=10, scale_hi=16, ef_lo=8,
                                          ef_hi=32, wmin=1, wmax=255)
    R, C, E = g.arrays()[:3]
    fut = _POOL.submit(O.dijkstra_ranges, R, C, E, voff, src)
    d = K.DeviceGraph(g)
    d.set_model(S.load_model())
    yield g, voff, src, d, fut
    d.close()


def test_graphs1000_sizes(graphs1000):
    g, voff, src, d, _ = graphs1000
    assert len(src) == 1000 and len(voff) == 1001
    sizes = np.diff(voff.astype(np.int64))
    assert sizes.min() >= 1 << 10 and sizes.max() <= 1 << 16
    assert d.m > 200_000_000


@pytest.mark.parametrize("sched", ["plain", "learned"])
def test_graphs1000_against_per_graph_dijkstra(graphs1000, sched):
    g, voff, src, d, fut = graphs1000
    shape = {}
    if sched == "learned":
        shape, _, _ = d.autotune("sssp", src, sched="learned")
    case(f"graphs1000 sssp sched={sched} shape={shape}")
    d.init_sources("sssp", src)
    info = d.run("sssp", sched=sched, **shape)
    got = d.values()
    exp = fut.result()
    bad = np.flatnonzero(got != exp)
    if len(bad):
        gi = np.searchsorted(voff, bad[:5], side="right") - 1
        pytest.fail(f"{len(bad)} vertices differ (graphs {sorted(set(gi.tolist()))})")
    assert info["iterations"] >= 2
