"""Host graph construction: the edge-list parser and CSR builder.

Three implementations must agree array for array: the reference's own
kites/graph/graph_raw.h compiled into oracle/_ref (when the reference
checkout is present), the oracle's Python restatement
(oracle/graph_raw_oracle.py) and the product's C++ parser
(kites_amd/csrc/kites/graph.hpp, through the C ABI).  Generators are checked
for the properties their configurations fix.
"""
import gzip
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from oracle import graph_raw_oracle as gro

K = pytest.importorskip("kites_amd")
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "graph_raw_driver")

CRAFTED = (
    "# comment line\r\n"
    "% another comment\n"
    "\n"
    "5 3\r\n"
    "1\t2\t7\n"          # tab separated, weight column
    "2 1 9 extra tokens\n"
    "5 3\n"              # duplicate edge (kept)
    "0 0\n"              # self loop (kept by the parser)
    "  4 1\n"            # leading space: skipped (first char not a digit)
    "3 4 \r\n"           # trailing space then CR: one extra token "\r"
    "130 2\n"            # pushes the vertex count past one padding block
    "7\n"                # single token: skipped
    "30 6 12\n"
    "6 30\n"
)

CONFIGS = [  # vpg, s_d, directed, complete
    (1, 1, 1, 1), (2, 1, 1, 1), (4, 0, 1, 1), (64, 1, 1, 1), (1, 1, 0, 0), (1, 0, 0, 1),
]


def _read_ref(path):
    b = open(path, "rb").read()
    off = 0

    def u64():
        nonlocal off
        x = int.from_bytes(b[off:off + 8], "little")
        off += 8
        return x

    def arr():
        nonlocal off
        k = u64()
        a = np.frombuffer(b[off:off + 4 * k], np.uint32).copy()
        off += 4 * k
        return a

    n, m, t = u64(), u64(), u64()
    V, R, C, E, B, dR, dC = [arr() for _ in range(7)]
    return dict(n=n, m=m, total_out=t, V=V, R=R, C=C, E=E, frontier=B, dir_R=dR, dir_C=dC)


def _oracle(text, cfg, source=30):
    vpg, sd, dr, cp = cfg
    return gro.build(text, gro.example_callback(source), vpg, bool(sd), bool(dr), bool(cp))


def _product(path, cfg, source=30):
    vpg, sd, dr, cp = cfg
    h = K.HostGraph.load_edgelist(path, vpg=vpg, s_d=sd, directed=dr, complete=cp, source=source)
    R, C, E, *_ = h.arrays()
    V0, F, dR, dC = h.load_arrays()
    return dict(n=h.n, m=h.m, V=V0, R=R, C=C, E=E, frontier=F, dir_R=dR, dir_C=dC)


def _same(a, b, keys=("n", "m", "V", "R", "C", "E", "frontier", "dir_R", "dir_C")):
    for k in keys:
        x, y = a[k], b[k]
        if isinstance(x, np.ndarray) or isinstance(y, np.ndarray):
            assert np.array_equal(np.asarray(x, np.uint32), np.asarray(y, np.uint32)), k
        else:
            assert x == y, k


@pytest.fixture(scope="module")
def wiki(tmp_path_factory):
    p = tmp_path_factory.mktemp("wiki") / "wiki.txt"
    p.write_bytes(gzip.open(os.path.join(GOLDEN, "wikivote.txt.gz")).read())
    return p


@pytest.fixture(scope="module")
def crafted(tmp_path_factory):
    p = tmp_path_factory.mktemp("crafted") / "g.txt"
    p.write_bytes(CRAFTED.encode())
    return p


def test_wikivote_counts(wiki):
    g = _oracle(wiki.read_bytes().decode(), (1, 1, 1, 1))
    assert g["m"] == 103689 and g["n"] == 8320
    ids = np.unique(np.concatenate([g["C"], np.repeat(np.arange(g["n"]), np.diff(g["R"]))]))
    assert len(ids) == 7115 and ids.max() == 8297
    outdeg = np.diff(g["dir_R"])
    assert outdeg[30] == 5 and np.diff(g["R"])[30] == 23
    assert int(np.argmax(outdeg)) == 2565 and outdeg[2565] == 893


@pytest.mark.parametrize("cfg", CONFIGS)
def test_product_parser_matches_oracle(wiki, crafted, cfg):
    for path in (crafted, wiki):
        o = _oracle(path.read_bytes().decode(), cfg)
        p = _product(path, cfg)
        keys = ["n", "m", "V", "R", "C", "E", "frontier"]
        if cfg[2]:
            keys += ["dir_R", "dir_C"]
        _same(o, p, keys)


@pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="oracle/_ref not built (reference absent)")
@pytest.mark.parametrize("cfg", CONFIGS)
def test_oracle_parser_matches_reference_build(wiki, crafted, cfg, tmp_path):
    for path in (crafted, wiki):
        out = tmp_path / "ref.bin"
        subprocess.run([REF_DRIVER, str(path), str(out)] + [str(x) for x in cfg] + ["30"],
                       check=True)
        r = _read_ref(str(out))
        o = _oracle(path.read_bytes().decode(), cfg)
        assert r["n"] == o["n"] and r["m"] == o["m"] and r["total_out"] == o["total_out"]
        for k in ("V", "R", "C", "E", "frontier", "dir_C"):
            assert np.array_equal(r[k], o[k]), k
        if cfg[2]:
            assert np.array_equal(r["dir_R"], o["dir_R"])


def test_reference_drops_undirected_frontier(crafted):
    # graph_csr.h:100-108 copies frontier bits only inside the directed-only
    # out-list loop; the restatement records both behaviours.
    o = _oracle(crafted.read_bytes().decode(), (1, 1, 0, 0))
    assert o["frontier"].any() and not o["frontier_as_reference"].any()


def test_missing_file_raises():
    with pytest.raises(K.KitesError) as ei:
        K.HostGraph.load_edgelist("/nonexistent/graph.txt")
    assert "Failed to open specified file" in str(ei.value)


def _check_csr(h, wmin, wmax, simple=True):
    R, C, E, oR, oC, oE = h.arrays()
    n = len(R) - 1
    assert R[0] == 0 and np.all(np.diff(R.astype(np.int64)) >= 0) and R[-1] == len(C)
    assert len(C) == 0 or C.max() < n
    assert np.all((E >= wmin) & (E <= wmax))
    dst = np.repeat(np.arange(n, dtype=np.int64), np.diff(R).astype(np.int64))
    key = dst * n + C.astype(np.int64)
    if simple:
        assert np.all(C.astype(np.int64) != dst)
        assert np.all(np.diff(key) > 0)     # sorted by (dst, src), no duplicates
    # out-CSR is the transpose with the same weights
    osrc = np.repeat(np.arange(n, dtype=np.int64), np.diff(oR).astype(np.int64))
    a = np.lexsort((C, dst))
    b = np.lexsort((osrc, oC))
    assert np.array_equal(C[a], osrc[b]) and np.array_equal(dst[a], oC[b])
    assert np.array_equal(E[a], oE[b])
    return R, C, E, oR, oC, oE


def test_rmat_properties_and_determinism():
    h1 = K.HostGraph.rmat(12, 16, 5, wmin=1, wmax=255)
    h2 = K.HostGraph.rmat(12, 16, 5, wmin=1, wmax=255)
    h3 = K.HostGraph.rmat(12, 16, 6, wmin=1, wmax=255)
    R1, C1, E1, oR1, *_ = _check_csr(h1, 1, 255)
    R2, C2, E2, *_ = h2.arrays()
    R3, C3, *_ = h3.arrays()
    assert np.array_equal(R1, R2) and np.array_equal(C1, C2) and np.array_equal(E1, E2)
    assert not (np.array_equal(R1, R3) and np.array_equal(C1, C3))
    n = 1 << 12
    assert len(R1) == n + 1 and 0.7 * 16 * n < len(C1) <= 16 * n
    # a=.57 skews both endpoints towards low ids: vertex 0 is the biggest hub
    outdeg = np.diff(oR1)
    assert int(np.argmax(outdeg)) == 0 and outdeg[0] > 20 * outdeg.mean()


def test_rmat_exact_edges_c1():
    h = K.HostGraph.rmat(13, 0, 1, wmin=1, wmax=64, n_limit=7115, exact_edges=103689)
    n, n_real, m = h.info()
    assert (n, n_real, m) == (7168, 7115, 103689)
    R, C, *_ = _check_csr(h, 1, 64)
    assert C.max() < 7115 and np.all(np.diff(R)[7115:] == 0)


def test_er_and_grid():
    h = K.HostGraph.erdos_renyi(3000, 48000, 3, wmin=1, wmax=9)
    R, C, E, oR, *_ = _check_csr(h, 1, 9)
    assert len(R) - 1 == 3072 and 46000 < len(C) <= 48000
    rows, cols = 13, 11
    g = K.HostGraph.grid(rows, cols, 4, wmin=1, wmax=100)
    R, C, E, *_ = _check_csr(g, 1, 100)
    assert len(C) == 2 * rows * (cols - 1) + 2 * cols * (rows - 1)
    deg = np.diff(R)[: rows * cols]
    assert deg.min() == 2 and deg.max() == 4


def test_symmetrize():
    h = K.HostGraph.rmat(10, 8, 2, wmin=1, wmax=5)
    s = h.symmetrize()
    R, C, *_ = _check_csr(s, 1, 5)
    n = len(R) - 1
    dst = np.repeat(np.arange(n, dtype=np.int64), np.diff(R).astype(np.int64))
    fwd = set(zip(C.tolist(), dst.tolist()))
    assert all((d, c) in fwd for c, d in fwd)


def test_rmat_batch_union():
    g, voff, src = K.HostGraph.rmat_batch(40, 9, scale_lo=6, scale_hi=10, ef_lo=4, ef_hi=12,
                                          wmin=1, wmax=255)
    g2, voff2, src2 = K.HostGraph.rmat_batch(40, 9, scale_lo=6, scale_hi=10, ef_lo=4, ef_hi=12,
                                             wmin=1, wmax=255)
    R, C, E, oR, oC, oE = _check_csr(g, 1, 255)
    assert np.array_equal(voff, voff2) and np.array_equal(src, src2)
    assert np.array_equal(C, g2.arrays()[1])
    assert voff[0] == 0 and voff[-1] == len(R) - 1
    sizes = np.diff(voff)
    assert set(np.unique(sizes)) <= {64, 128, 256, 512, 1024} and len(np.unique(sizes)) > 1
    outdeg = np.diff(oR.astype(np.int64))
    for i in range(40):
        lo, hi = int(voff[i]), int(voff[i + 1])
        # every in-edge of a vertex of graph i comes from graph i
        seg = C[R[lo]:R[hi]]
        assert len(seg) == 0 or (seg.min() >= lo and seg.max() < hi)
        # the source is graph i's max-out-degree vertex, lowest id on ties
        assert src[i] == lo + int(np.argmax(outdeg[lo:hi]))


@pytest.mark.parametrize("n,m,seed", [(300, 3000, 1), (1, 5, 2), (5000, 40000, 5)])
def test_host_edge_build_matches_oracle(n, m, seed):
    """HostGraph.from_edges (the reference for the GPU build at scale) equals
    the oracle's restatement of graph_raw.h on lists with duplicates,
    self-loops and edgeless vertices."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    src[: max(1, m // 10)] = src[0]
    w = rng.integers(0, 1000, m).astype(np.uint32)
    nid = int(max(src.max(), dst.max())) + 1
    text = "".join(f"{a} {b} {c}\n" for a, b, c in zip(src, dst, w))
    ref = gro.build(text, gro.example_callback(0), 1, True, True, True)
    R, C, E, oR, oC, oE = K.HostGraph.from_edges(nid, src, dst, w).arrays()
    assert np.array_equal(R, ref["R"]) and np.array_equal(C, ref["C"])
    assert np.array_equal(E, ref["E"])
    assert np.array_equal(oR, ref["dir_R"]) and np.array_equal(oC, ref["dir_C"])


def test_host_edge_build_dedup_and_range():
    src = np.array([3, 1, 3, 2, 3, 5], np.uint32)
    dst = np.array([4, 1, 4, 4, 4, 0], np.uint32)
    w = np.array([7, 9, 8, 6, 5, 1], np.uint32)
    R, C, E, oR, oC, oE = K.HostGraph.from_edges(6, src, dst, w, dedup=True).arrays()
    # self-loop 1->1 dropped; 3->4 kept once with the first weight
    assert len(C) == 3 and list(C[R[4]:R[5]]) == [2, 3] and list(E[R[4]:R[5]]) == [6, 7]
    assert list(C[R[0]:R[1]]) == [5]
    with pytest.raises(K.KitesError):
        K.HostGraph.from_edges(5, src, dst, w)


# ---- prepared-graph files --------------------------------------------------

@pytest.mark.parametrize("cfg", CONFIGS)
def test_text_dump_matches_reference_format(wiki, crafted, cfg, tmp_path):
    """save_text writes, byte for byte, what the reference's save_to_file
    writes for the same parsed graph (restated in the oracle)."""
    for path in (crafted, wiki):
        vpg, sd, dr, cp = cfg
        h = K.HostGraph.load_edgelist(path, vpg=vpg, s_d=sd, directed=dr, complete=cp, source=30)
        out = tmp_path / "dump.txt"
        h.save_text(out)
        exp = gro.dump_text(_oracle(path.read_bytes().decode(), cfg), bool(dr))
        assert out.read_bytes() == exp.encode()


@pytest.mark.parametrize("cfg", [c for c in CONFIGS if c[0] == 1 and c[2]])
def test_text_roundtrip(wiki, crafted, cfg, tmp_path):
    """graph_csr(file) semantics: a directed one-vertex-per-group dump reads
    back to the same arrays, values and frontier, and writes the same bytes."""
    for path in (crafted, wiki):
        vpg, sd, dr, cp = cfg
        h = K.HostGraph.load_edgelist(path, vpg=vpg, s_d=sd, directed=dr, complete=cp, source=30)
        a = tmp_path / "a.txt"
        b = tmp_path / "b.txt"
        h.save_text(a)
        g = K.HostGraph.load_text(a)
        for x, y in zip(g.arrays(), h.arrays()):
            assert np.array_equal(x, y)
        V0, F, dR, dC = h.load_arrays()
        V1, F1, dR1, dC1 = g.load_arrays()
        assert np.array_equal(V0, V1) and np.array_equal(F, F1)
        assert np.array_equal(dR, dR1) and np.array_equal(dC, dC1)
        g.save_text(b)
        assert a.read_bytes() == b.read_bytes()


def test_text_reader_rejects(tmp_path, wiki):
    h = K.HostGraph.load_edgelist(wiki)
    p = tmp_path / "d.txt"
    h.save_text(p)
    lines = p.read_text().split("\n")
    (tmp_path / "trunc.txt").write_text("\n".join(lines[: len(lines) // 2]))
    with pytest.raises(K.KitesError):
        K.HostGraph.load_text(tmp_path / "trunc.txt")
    bad = lines[:]
    bad[5] = "x"
    (tmp_path / "bad.txt").write_text("\n".join(bad))
    with pytest.raises(K.KitesError):
        K.HostGraph.load_text(tmp_path / "bad.txt")
    with pytest.raises(K.KitesError):
        K.HostGraph.load_text(tmp_path / "missing.txt")


def test_binary_roundtrip(wiki, tmp_path):
    for h in (K.HostGraph.load_edgelist(wiki), K.HostGraph.rmat(14, 8, 3, wmin=1, wmax=99),
              K.HostGraph.load_edgelist(wiki, vpg=4)):
        p = tmp_path / "g.bin"
        h.save_binary(p)
        g = K.HostGraph.load_binary(p)
        assert (g.n, g.m) == (h.n, h.m)
        for x, y in zip(g.arrays(), h.arrays()):
            assert np.array_equal(x, y)
        for x, y in zip(g.load_arrays(), h.load_arrays()):
            assert np.array_equal(x, y)
    raw = p.read_bytes()
    (tmp_path / "t.bin").write_bytes(raw[: len(raw) // 2])
    with pytest.raises(K.KitesError):
        K.HostGraph.load_binary(tmp_path / "t.bin")
    (tmp_path / "m.bin").write_bytes(b"NOTAGRPH" + raw[8:])
    with pytest.raises(K.KitesError):
        K.HostGraph.load_binary(tmp_path / "m.bin")


def test_binary_reader_rejects_malformed_headers_and_group_lists(wiki, tmp_path):
    """A binary graph file is untrusted input: every size in its header and
    every entry of the vertex-group lists (which the iteration kernel walks
    when vertex-per-group > 1) is checked before the graph can be used."""
    import struct
    h = K.HostGraph.load_edgelist(wiki, vpg=4)
    p = tmp_path / "g.bin"
    h.save_binary(p)
    raw = bytearray(p.read_bytes())
    n, n_real, vpg, directed = struct.unpack_from("<4I", raw, 8)
    m, nV0, nfront, ndirR, ndirC = struct.unpack_from("<5Q", raw, 24)
    ngroups = (n + vpg - 1) // vpg
    assert vpg == 4 and 1 < ndirR <= ngroups + 1 and ndirC > 0
    hdr = 64
    off_dirR = hdr + 4 * (2 * (n + 1) + 4 * m + nV0) + nfront
    off_dirC = off_dirR + 4 * ndirR
    assert off_dirC + 4 * ndirC == len(raw)

    def rejects(mutate, name):
        b = bytearray(raw)
        mutate(b)
        q = tmp_path / (name + ".bin")
        q.write_bytes(bytes(b))
        with pytest.raises(K.KitesError):
            K.HostGraph.load_binary(q)

    rejects(lambda b: struct.pack_into("<I", b, 16, 3), "vpg3")                 # not a power of two
    rejects(lambda b: struct.pack_into("<I", b, 16, 0), "vpg0")
    rejects(lambda b: struct.pack_into("<I", b, 12, n + 1), "nreal")           # n_real > n
    rejects(lambda b: struct.pack_into("<Q", b, 24, 1 << 40), "mhuge")         # m beyond 32 bits
    rejects(lambda b: struct.pack_into("<Q", b, 24, m + 10), "mtrunc")         # more than the file holds
    rejects(lambda b: struct.pack_into("<Q", b, 32, n + 1), "nv0")             # values != n
    rejects(lambda b: struct.pack_into("<Q", b, 40, n + 1), "nfront")          # frontier > n
    rejects(lambda b: struct.pack_into("<Q", b, 48, ngroups + 2), "ndirr")     # more groups than n allows
    rejects(lambda b: struct.pack_into("<Q", b, 48, 0), "ndirr0")
    rejects(lambda b: struct.pack_into("<I", b, len(b) - 4, ngroups), "gid")   # group id == groups
    rejects(lambda b: struct.pack_into("<I", b, len(b) - 4, 0xFFFFFFFF), "gidmax")
    rejects(lambda b: struct.pack_into("<I", b, off_dirR, 1), "dir0")          # dir_R[0] != 0
    rejects(lambda b: struct.pack_into("<I", b, off_dirC - 4, ndirC - 1), "dirend")
    first = struct.unpack_from("<I", raw, off_dirR + 4)[0]
    rejects(lambda b: struct.pack_into("<I", b, off_dirR + 8, 0) if first > 0 else
            struct.pack_into("<I", b, off_dirR + 4, ndirC + 5), "dirdec")       # offsets decrease
    # the untouched file still loads
    g = K.HostGraph.load_binary(p)
    assert (g.n, g.m) == (h.n, h.m)
