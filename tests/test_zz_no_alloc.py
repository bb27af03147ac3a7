"""GPU test of kites_sssp_u32_noalloc, the entry point that allocates
nothing: two calls in a row (the second on a smaller graph, so the fixed
device array still holds the first call's larger state) must give exactly
what kites_sssp_u32 gives on the same graphs, which is the oracle's
fixpoint; sizes beyond the array's limits are refused.
"""
import numpy as np
import pytest

from conftest import TINY_GRAPHS, case, tiny_graph

pytestmark = pytest.mark.gpu

K = pytest.importorskip("kites_amd")
from kites_amd._lib import KITES_ERR_ARG  # noqa: E402
from oracle import liboracle as O  # noqa: E402


def _maxdeg_source(h):
    oR = h.arrays()[3]
    return int(np.argmax(np.diff(oR)))


def _oracle_sssp(h, src):
    R, C, E = h.arrays()[:3]
    V0, _ = O.sssp_init(len(R) - 1, src, R, C)
    V, _ = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
    return V


@pytest.mark.parametrize("sched", ["plain", "classic", "learned"])
def test_noalloc_two_calls_equal_the_allocating_entry_point(sched):
    big = K.HostGraph.rmat(16, 16, 71, wmin=1, wmax=255)
    small = K.HostGraph.rmat(11, 8, 72, wmin=1, wmax=64)
    model = None
    if sched == "learned":
        from kites_amd import scheduler
        model = scheduler.load_model()
    for h in (big, small):
        src = _maxdeg_source(h)
        d = K.DeviceGraph(h)
        if model is not None:
            d.set_model(model)
        case(f"noalloc sssp n={d.n} m={d.m} src={src} sched={sched}")
        got, info = d.sssp_noalloc(src, sched=sched)
        exp, _ = d.solve("sssp", src, sched=sched)
        assert np.array_equal(got, exp), f"{int((got != exp).sum())} vertices differ"
        assert info["iterations"] >= 1
        if d.n <= (1 << 12):
            assert np.array_equal(got, _oracle_sssp(h, src))
        # the same graph again: the array's leftover state is reset
        again, _ = d.sssp_noalloc(src, sched=sched)
        assert np.array_equal(again, exp)
        d.close()


def test_noalloc_refuses_sizes_beyond_the_limits():
    h = K.HostGraph.rmat(10, 8, 73)
    d = K.DeviceGraph(h)
    src = _maxdeg_source(h)
    case("noalloc limits: max_iterations, split_pull, keep_stats")
    for kw in (dict(max_iterations=(1 << 16) + 1), dict(split_pull=True), dict(keep_stats=True)):
        with pytest.raises(K.KitesError) as e:
            d.sssp_noalloc(src, **kw)
        assert e.value.code == KITES_ERR_ARG
    d.close()
    n = (1 << 22) + 128            # one 128-vertex block above the limit, no edges
    R = np.zeros(n + 1, np.uint32)
    C = np.zeros(0, np.uint32)
    hb = K.HostGraph.from_csr(R, C, C)
    db = K.DeviceGraph(hb)
    case(f"noalloc limits: n={db.n}")
    with pytest.raises(K.KitesError) as e:
        db.sssp_noalloc(0)
    assert e.value.code == KITES_ERR_ARG
    assert "4,194,304" in str(e.value)
    db.close()


@pytest.mark.parametrize("name", TINY_GRAPHS)
def test_noalloc_tiny_graphs(name):
    """The tiny graphs (no edges, a lone self-loop, self-loops and parallel
    edges over 70 vertices) right after a larger run, so the array holds
    more state than the tiny run writes: the oracle's fixpoint from every
    vertex's run (plain) and from the first and last (learned)."""
    from kites_amd import scheduler
    big = K.HostGraph.rmat(12, 8, 74, wmin=1, wmax=64)
    db = K.DeviceGraph(big)
    case(f"noalloc sssp rmat12 before {name}")
    got, _ = db.sssp_noalloc(_maxdeg_source(big), sched="plain")
    assert np.array_equal(got, _oracle_sssp(big, _maxdeg_source(big)))
    db.close()
    h = tiny_graph(name)
    d = K.DeviceGraph(h)
    d.set_model(scheduler.load_model())
    for src in range(d.n):
        case(f"noalloc sssp {name} from {src}")
        got, _ = d.sssp_noalloc(src, sched="plain")
        assert np.array_equal(got, _oracle_sssp(h, src)), src
    for src in sorted({0, d.n - 1}):
        got, _ = d.sssp_noalloc(src, sched="learned")
        assert np.array_equal(got, _oracle_sssp(h, src)), src
    d.close()
