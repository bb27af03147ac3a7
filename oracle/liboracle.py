"""ctypes binding of the C oracle (``kites_oracle.c``).  Test infrastructure only."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libkites_oracle.so")

SSSP_SAT, BFS, CC, SSSP_WRAP, WIDEST = 0, 1, 2, 3, 4
INF = 0xFFFFFFFF

_lib = None


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        L.oracle_fixpoint_u32.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, u32p,
                                          ctypes.c_int, ctypes.c_uint32]
        L.oracle_fixpoint_u32.restype = ctypes.c_uint32
        L.oracle_frontier_rounds_u32.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, u32p,
                                                 u8p, ctypes.c_int, ctypes.c_uint32]
        L.oracle_frontier_rounds_u32.restype = ctypes.c_uint32
        L.oracle_dijkstra_u32.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, ctypes.c_uint32, u32p]
        L.oracle_widest_u32.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, ctypes.c_uint32, u32p]
        L.oracle_bfs_u32.argtypes = [ctypes.c_uint32, u32p, u32p, ctypes.c_uint32, u32p]
        L.oracle_cc_min_label.argtypes = [ctypes.c_uint32, u32p, u32p, u32p]
        L.oracle_frontier_omp_u32.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, u32p, u8p,
                                              ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_int)]
        L.oracle_frontier_omp_u32.restype = ctypes.c_uint32
        L.oracle_frontier_omp_csr_u32.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, u32p, u32p,
                                                  u32p, u32p, u8p, ctypes.c_int, ctypes.c_int,
                                                  ctypes.POINTER(ctypes.c_int)]
        L.oracle_frontier_omp_csr_u32.restype = ctypes.c_uint32
        L.oracle_check_cols_u32.argtypes = [ctypes.c_uint32, u32p, u32p, u32p, ctypes.c_uint32,
                                            u32p, u32p, ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_check_cols_u32.restype = ctypes.c_uint64
        L.oracle_dijkstra_ranges_u32.argtypes = [ctypes.c_uint32, u32p, u32p, u32p,
                                                 ctypes.c_uint32, u32p, u32p, ctypes.c_int, u32p]
        L.oracle_dijkstra_ranges_u32.restype = ctypes.c_int
        L.oracle_affinity_cores.restype = ctypes.c_int
        f32p = ctypes.POINTER(ctypes.c_float)
        L.oracle_ordered_pagerank_f32.argtypes = [ctypes.c_uint32, u32p, u32p, f32p, u32p, u32p,
                                                  f32p, u8p, ctypes.c_float, ctypes.c_float,
                                                  ctypes.c_float, ctypes.c_uint32]
        L.oracle_ordered_pagerank_f32.restype = ctypes.c_uint32
        _lib = L
    return _lib


def _p(a, t=ctypes.c_uint32):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(t))


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def fixpoint(R, C, E, V0, algo, max_sweeps=0):
    """Gauss-Seidel sweep (sanity_check.cuh:21-47). Returns (V, sweeps)."""
    R, C = _u32(R), _u32(C)
    E = None if E is None else _u32(E)
    V = _u32(V0).copy()
    n = len(V)
    sweeps = lib().oracle_fixpoint_u32(n, _p(R), _p(C), _p(E), _p(V), algo, max_sweeps)
    return V, sweeps


def frontier_rounds(R, C, E, V0, frontier0, algo, max_rounds=0):
    R, C = _u32(R), _u32(C)
    E = None if E is None else _u32(E)
    V = _u32(V0).copy()
    F = np.ascontiguousarray(frontier0, dtype=np.uint8)
    rounds = lib().oracle_frontier_rounds_u32(len(V), _p(R), _p(C), _p(E), _p(V),
                                              _p(F, ctypes.c_uint8), algo, max_rounds)
    return V, rounds


def frontier_omp(R, C, E, V0, frontier0, algo, threads=0):
    R, C = _u32(R), _u32(C)
    E = None if E is None else _u32(E)
    V = _u32(V0).copy()
    F = np.ascontiguousarray(frontier0, dtype=np.uint8)
    used = ctypes.c_int(0)
    rounds = lib().oracle_frontier_omp_u32(len(V), _p(R), _p(C), _p(E), _p(V),
                                           _p(F, ctypes.c_uint8), algo, threads,
                                           ctypes.byref(used))
    return V, rounds, used.value


def frontier_omp_csr(R, C, E, oR, oC, oE, V0, frontier0, algo, threads=0):
    """frontier_omp over the out-CSR the caller already holds (no transpose
    inside the call).  oE None = unit weights."""
    R, C, oR, oC = _u32(R), _u32(C), _u32(oR), _u32(oC)
    E = None if E is None else _u32(E)
    oE = None if oE is None else _u32(oE)
    V = _u32(V0).copy()
    F = np.ascontiguousarray(frontier0, dtype=np.uint8)
    used = ctypes.c_int(0)
    rounds = lib().oracle_frontier_omp_csr_u32(len(V), _p(R), _p(C), _p(E), _p(oR), _p(oC),
                                               _p(oE), _p(V), _p(F, ctypes.c_uint8), algo,
                                               threads, ctypes.byref(used))
    return V, rounds, used.value


def check_cols(R, C, E, D, sources, threads=0):
    """Shortest-path certificate of the columns of D (n x K) from sources:
    returns (total, [bad sources, violated edges, untight entries]).  E None
    checks BFS levels (unit weights)."""
    R, C = _u32(R), _u32(C)
    E = None if E is None else _u32(E)
    D = np.ascontiguousarray(D, np.uint32)
    src = _u32(sources)
    n = len(R) - 1
    assert D.ndim == 2 and D.shape == (n, len(src)), (D.shape, n, len(src))
    bad = (ctypes.c_uint64 * 3)()
    tot = lib().oracle_check_cols_u32(n, _p(R), _p(C), _p(E), len(src), _p(D), _p(src),
                                      threads, bad)
    return int(tot), [int(x) for x in bad]


def dijkstra_ranges(R, C, E, voff, sources, threads=0):
    """Per-graph Dijkstra over a disjoint union (graph i = vertices
    [voff[i], voff[i+1]) from sources[i]), graphs in parallel."""
    R, C, E, voff, src = _u32(R), _u32(C), _u32(E), _u32(voff), _u32(sources)
    n = len(R) - 1
    d = np.full(n, INF, np.uint32)
    err = lib().oracle_dijkstra_ranges_u32(n, _p(R), _p(C), _p(E), len(src), _p(voff), _p(src),
                                           threads, _p(d))
    if err:
        raise ValueError("dijkstra_ranges: the graph is not a disjoint union of the ranges")
    return d


def dijkstra(R, C, E, src):
    R, C, E = _u32(R), _u32(C), _u32(E)
    n = len(R) - 1
    d = np.empty(n, np.uint32)
    lib().oracle_dijkstra_u32(n, _p(R), _p(C), _p(E), src, _p(d))
    return d


def widest(R, C, E, src):
    R, C, E = _u32(R), _u32(C), _u32(E)
    n = len(R) - 1
    d = np.empty(n, np.uint32)
    lib().oracle_widest_u32(n, _p(R), _p(C), _p(E), src, _p(d))
    return d


def bfs(R, C, src):
    R, C = _u32(R), _u32(C)
    n = len(R) - 1
    d = np.empty(n, np.uint32)
    lib().oracle_bfs_u32(n, _p(R), _p(C), src, _p(d))
    return d


def cc_min_label(R, C):
    R, C = _u32(R), _u32(C)
    n = len(R) - 1
    d = np.empty(n, np.uint32)
    lib().oracle_cc_min_label(n, _p(R), _p(C), _p(d))
    return d


def ordered_pagerank(R, C, W, oR, oC, V0, active0, base=0.15, damp=0.85, tol=1e-4,
                     max_iter=10000):
    """Ordered Jacobi PageRank in float32 with the engine's combine order
    (kites_oracle.c).  Returns (V, iterations); iterations 0 = no convergence."""
    R, C, oR, oC = _u32(R), _u32(C), _u32(oR), _u32(oC)
    W = np.ascontiguousarray(W, np.float32)
    V = np.ascontiguousarray(V0, np.float32).copy()
    a = np.ascontiguousarray(active0, np.uint8)
    f32p = ctypes.POINTER(ctypes.c_float)
    it = lib().oracle_ordered_pagerank_f32(len(R) - 1, _p(R), _p(C), W.ctypes.data_as(f32p),
                                           _p(oR), _p(oC), V.ctypes.data_as(f32p),
                                           a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                           base, damp, tol, max_iter)
    return V, it


def affinity_cores():
    return lib().oracle_affinity_cores()


# ---- standard initial states (same rules the engine's C-ABI uses) ----

def sssp_init(n, src, R, C):
    """V = INF except V[src]=0; frontier = out-neighbours of src (the
    reference callback's rule: vBitmap[dst] |= (src == source),
    examples/sssp_basic_single_device.cu:21-25)."""
    V = np.full(n, INF, np.uint32)
    V[src] = 0
    F = np.zeros(n, np.uint8)
    C = np.asarray(C)
    R = np.asarray(R)
    dst_of_edge = np.repeat(np.arange(n, dtype=np.uint32), np.diff(R).astype(np.int64))
    F[dst_of_edge[C == src]] = 1
    return V, F


def widest_init(n, src, R, C):
    V = np.zeros(n, np.uint32)
    V[src] = INF
    _, F = sssp_init(n, src, R, C)
    return V, F


def cc_init(n):
    return np.arange(n, dtype=np.uint32), np.ones(n, np.uint8)
