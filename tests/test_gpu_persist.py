"""The one-workgroup loop for small push frontiers (persist_kernel.hpp).

After a small push frontier the engine's next chunk starts with a launch of
one workgroup that runs every iteration whose queue fits it (up to 4,096
entries and 2,048 out-edges by default) before the step launches take over.  The
results must be the oracle's bit for bit, with the loop and without it
(options persist on and off), on graphs whose runs are long chains of small
frontiers (grids, paths, a caterpillar whose queue entries have very uneven
degrees), on R-MAT (a few small tail iterations after large ones), under the
direct-launch path, under the iteration limit and with near-far (far-pile
appends and splits inside the loop).  The statistics of a run must chain:
every iteration consumes the queue the previous one appended.
"""
import numpy as np
import pytest

from conftest import case

pytestmark = pytest.mark.gpu

K = pytest.importorskip("kites_amd")
from kites_amd import scheduler as S  # noqa: E402
from oracle import liboracle as O  # noqa: E402


def _caterpillar(spine, leaves_every, leaves, seed):
    """A path of `spine` vertices; every `leaves_every`-th spine vertex gets
    `leaves` leaf vertices; both directions of every edge, weights U[1,50]."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    s = np.arange(spine - 1, dtype=np.uint32)
    src += [s, s + 1]
    dst += [s + 1, s]
    nxt = spine
    for v in range(0, spine, leaves_every):
        lv = np.arange(nxt, nxt + leaves, dtype=np.uint32)
        nxt += leaves
        hub = np.full(leaves, v, np.uint32)
        src += [hub, lv]
        dst += [lv, hub]
    src = np.concatenate(src)
    dst = np.concatenate(dst)
    w = rng.integers(1, 51, size=len(src), dtype=np.uint32)
    return K.HostGraph.from_edges(nxt, src, dst, w)


GRAPHS = {
    "grid96x80": lambda: K.HostGraph.grid(96, 80, 7, wmin=1, wmax=100),
    "grid512": lambda: K.HostGraph.grid(512, 512, 11, wmin=1, wmax=100),
    "caterpillar": lambda: _caterpillar(6000, 40, 120, 5),
    "rmat16": lambda: K.HostGraph.rmat(16, 16, 16, wmin=1, wmax=255),
}
# graphs whose SSSP and BFS runs are long chains of small push frontiers
LONG = {"grid96x80", "caterpillar"}


@pytest.fixture(scope="module", params=sorted(GRAPHS))
def graph(request):
    name = request.param
    h = GRAPHS[name]()
    if name in ("rmat16",):
        hs = h.symmetrize()
    else:
        hs = h   # built with both directions of every edge
    d = K.DeviceGraph(h)
    d.set_model(S.load_model())
    ds = d if hs is h else K.DeviceGraph(hs)
    if ds is not d:
        ds.set_model(S.load_model())
    R, C, E = h.arrays()[:3]
    Rs, Cs = hs.arrays()[:2]
    exp = {"sssp": O.dijkstra(R, C, E, 0), "bfs": O.bfs(R, C, 0), "cc": O.cc_min_label(Rs, Cs)}
    yield name, d, ds, exp
    if ds is not d:
        ds.close()
    d.close()


def _chained(st):
    """Every iteration after the first consumes the previous one's queue."""
    for i in range(1, len(st)):
        if st[i]["in_count"] != st[i - 1]["count"]:
            return i
    return None


@pytest.mark.parametrize("sched", ["plain", "learned"])
@pytest.mark.parametrize("algo", ["sssp", "bfs", "cc"])
def test_persist_matches_oracle(graph, algo, sched):
    name, d, ds, exp = graph
    g = ds if algo == "cc" else d
    case(f"persist {name} {algo} {sched}")
    got, info = g.solve(algo, 0, sched=sched, keep_stats=True)
    st = g.stats()
    assert np.array_equal(got, exp[algo]), \
        f"{int((got != exp[algo]).sum())} vertices differ from the oracle"
    assert _chained(st) is None, f"iteration {_chained(st)} does not consume its predecessor's queue"
    nper = sum(1 for s in st if K.VARIANTS[s["variant"]] == "persist")
    off, info_off = g.solve(algo, 0, sched=sched, persist=False, keep_stats=True)
    st_off = g.stats()
    assert np.array_equal(off, got)
    assert not any(K.VARIANTS[s["variant"]] == "persist" for s in st_off)
    print(f"{name} {algo} {sched}: {info['iterations']} it ({nper} in the loop) "
          f"{info['ms']:.3f} ms; without {info_off['iterations']} it {info_off['ms']:.3f} ms",
          flush=True)
    if name in LONG and algo != "cc":
        assert nper > info["iterations"] // 2


@pytest.mark.parametrize("sched", ["plain", "learned"])
@pytest.mark.parametrize("delta", [10, 50, 200])
def test_persist_near_far(graph, delta, sched):
    """Near-far SSSP and BFS: the loop appends far entries to the pile and
    runs splits; the distances are the oracle's with the loop and without."""
    name, d, ds, exp = graph
    for algo in ("sssp", "bfs"):
        case(f"persist near-far {name} {algo} delta={delta} {sched}")
        got, info = d.solve(algo, 0, sched=sched, delta=delta, keep_stats=True)
        st = d.stats()
        assert np.array_equal(got, exp[algo]), \
            f"{algo}: {int((got != exp[algo]).sum())} vertices differ from the oracle"
        names = [K.VARIANTS[s["variant"]] for s in st]
        off, info_off = d.solve(algo, 0, sched=sched, delta=delta, persist=False)
        assert np.array_equal(off, got)
        print(f"{name} {algo} delta={delta} {sched}: {info['iterations']} it "
              f"({names.count('persist')} loop pushes, {names.count('split')} splits) "
              f"{info['ms']:.3f} ms; without the loop {info_off['iterations']} it "
              f"{info_off['ms']:.3f} ms", flush=True)
        if name in LONG and algo == "sssp" and delta <= 50:
            assert names.count("persist") > 0 and names.count("split") > 0


@pytest.mark.parametrize("name", ["grid96x80", "caterpillar"])
def test_persist_direct_launches(name):
    """No hipGraph: the loop is launched directly before each small chunk."""
    h = GRAPHS[name]()
    R, C, E = h.arrays()[:3]
    exp = O.dijkstra(R, C, E, 0)
    d = K.DeviceGraph(h)
    try:
        for chunk in (1, 3, 16):
            for delta in (0, 30):
                case(f"persist direct {name} chunk={chunk} delta={delta}")
                got, _ = d.solve("sssp", 0, sched="plain", use_graph=False, chunk=chunk,
                                 delta=delta)
                assert np.array_equal(got, exp), f"chunk {chunk} delta {delta}"
    finally:
        d.close()


def test_persist_iteration_limit():
    """A run cut by max_iterations fails the same way with and without the loop."""
    h = K.HostGraph.grid(96, 80, 7, wmin=1, wmax=100)
    d = K.DeviceGraph(h)
    try:
        for persist in (True, False):
            case(f"persist limit persist={persist}")
            with pytest.raises(K.KitesError):
                d.solve("bfs", 0, sched="plain", max_iterations=60, persist=persist)
        # and a limit the run fits in
        got, info = d.solve("bfs", 0, sched="plain", max_iterations=400)
        R, C = h.arrays()[:2]
        assert np.array_equal(got, O.bfs(R, C, 0))
        assert info["iterations"] <= 400
    finally:
        d.close()


def test_persist_repeated_runs_other_sources():
    """Runs on one graph replay cached chunks; every source gives the oracle's
    distances."""
    h = GRAPHS["caterpillar"]()
    R, C, E = h.arrays()[:3]
    d = K.DeviceGraph(h)
    try:
        for src in (0, 2999, 5999, 6100, 0):
            case(f"persist repeated src={src}")
            got, _ = d.solve("sssp", src, sched="plain")
            assert np.array_equal(got, O.dijkstra(R, C, E, src)), f"source {src}"
    finally:
        d.close()


def test_persist_noalloc_entry_point():
    """kites_sssp_u32_noalloc launches the loop directly too."""
    h = GRAPHS["grid512"]()
    R, C, E = h.arrays()[:3]
    d = K.DeviceGraph(h)
    try:
        case("persist noalloc grid512")
        got, _ = d.sssp_noalloc(0, sched="plain")
        assert np.array_equal(got, O.dijkstra(R, C, E, 0))
    finally:
        d.close()
