"""Settled runs: BFS pulls that skip reached vertices and stop at the first
frontier in-neighbour (options::settled_below, set by the C ABI for BFS from
a unit start).

Results must equal the oracle's BFS with every variant and schedule, from
one source and from many; a BFS started from an arbitrary state
(kites_graph_set_state) is not settled and must equal the oracle's fixpoint
from that state; and the settled pulls must read fewer in-edges than the
graph has (the early exit is what makes them fast).
"""
import numpy as np
import pytest

from conftest import case

pytestmark = pytest.mark.gpu

K = pytest.importorskip("kites_amd")
from kites_amd import scheduler as S  # noqa: E402
from oracle import liboracle as O  # noqa: E402


@pytest.fixture(scope="module")
def rmat18():
    h = K.HostGraph.rmat(18, 16, 5, wmin=1, wmax=255)
    d = K.DeviceGraph(h)
    d.set_model(S.load_model())
    yield h, d
    d.close()


def _src(h):
    oR = h.arrays()[3]
    return int(np.argmax(np.diff(oR.astype(np.int64))))


@pytest.mark.parametrize("kw", [dict(sched="plain"), dict(sched="plain", alpha=2.0),
                                dict(sched="learned"), dict(sched="plain", gran=1),
                                dict(sched="plain", delta=3)] +
                         [dict(sched="forced", forced_variant=v)
                          for v in ("pull_t", "pull_w", "pull_b", "pull_lb")],
                         ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_settled_bfs_matches_oracle(rmat18, kw):
    h, d = rmat18
    R, C = h.arrays()[:2]
    src = _src(h)
    case(f"settled bfs rmat18 {kw}")
    got, _ = d.solve("bfs", src, **kw)
    exp = O.bfs(R, C, src)
    assert np.array_equal(got, exp), f"{int((got != exp).sum())} vertices differ"


def test_settled_pulls_stop_early(rmat18):
    """The plain schedule's BFS pulls (thread per vertex) read far fewer
    in-edges than m, while SSSP pulls read them all."""
    h, d = rmat18
    src = _src(h)
    case("settled bfs work")
    d.solve("bfs", src, sched="plain", keep_stats=True)
    st = d.stats()
    pulls = [s for s in st if K.VARIANTS[s["variant"]] == "pull_t"]
    assert pulls, "the plain schedule pulls on R-MAT BFS"
    for s in pulls:
        assert s["work_edges"] < d.m // 2, s
    case("sssp pull work")
    d.solve("sssp", src, sched="forced", forced_variant="pull_t", keep_stats=True)
    st = d.stats()
    assert all(s["work_edges"] == d.m for s in st[1:] if K.VARIANTS[s["variant"]] == "pull_t")


def test_multi_source_bfs_settled(rmat18):
    h, d = rmat18
    R, C = h.arrays()[:2]
    oR = h.arrays()[3]
    cand = np.flatnonzero(np.diff(oR.astype(np.int64)) > 0)
    srcs = cand[:: max(1, len(cand) // 5)][:5]
    # multi-source BFS levels: min over sources of each source's levels
    exp = np.min(np.stack([O.bfs(R, C, int(s)) for s in srcs]), axis=0)
    for sched in ("plain", "learned"):
        case(f"settled multi-source bfs {sched}")
        d.init_sources("bfs", srcs)
        d.run("bfs", sched=sched)
        assert np.array_equal(d.values(), exp)


def test_bfs_from_arbitrary_state_is_not_settled(rmat18):
    """From a state that is not a unit start, BFS values are not final when
    first written: the run must give the oracle's fixpoint from that state."""
    h, d = rmat18
    R, C, E = h.arrays()[:3]
    n = len(R) - 1
    rng = np.random.default_rng(11)
    V0 = np.full(n, K.INF, np.uint32)
    seeds = rng.choice(n, 50, replace=False)
    V0[seeds] = rng.integers(0, 40, 50).astype(np.uint32)
    F0 = np.ones(n, np.uint8)
    exp, _ = O.fixpoint(R, C, E, V0, O.BFS)
    for kw in (dict(sched="plain"), dict(sched="forced", forced_variant="pull_t")):
        case(f"bfs from set_state {kw}")
        d.set_state(V0, F0)
        d.run("bfs", **kw)
        assert np.array_equal(d.values(), exp)
