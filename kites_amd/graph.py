"""Python mirror of the engine's pipeline over the C ABI.

``HostGraph`` is the host CSR graph (generators, edge-list loading as the
reference's ``graph_raw`` does, ``kites/graph/graph_raw.h:58-190``).
``DeviceGraph`` is the counterpart of ``make_graph_csr_for_device``
(``kites/graph/graph_csr_device.cuh:146-152``) and ``DeviceGraph.run`` of
``kites::process<launch::sync>`` (``kites/device/compute.cuh:29-137``) with
the functor sets the library instantiates (``sssp``, ``bfs``, ``cc``,
``sssp_wrap`` = the reference example's lambdas, ``widest``).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import ALGO, INF, SCHED, VARIANTS, check, lib


def _arr(ptr, count, dtype=np.uint32):
    if count == 0 or not ptr:          # an empty array has no data pointer
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).astype(dtype, copy=True)


class HostGraph:
    """Host-side CSR graph owned by the native library."""

    def __init__(self, handle):
        self._h = ctypes.c_void_p(handle) if not isinstance(handle, ctypes.c_void_p) else handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().kites_hgraph_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    @staticmethod
    def _make(fn, *args):
        out = ctypes.c_void_p()
        check(fn(*args, ctypes.byref(out)))
        return HostGraph(out)

    @classmethod
    def rmat(cls, scale, edge_factor, seed, wmin=1, wmax=255, a=0.57, b=0.19, c=0.19,
             n_limit=0, exact_edges=0, threads=0):
        return cls._make(lib().kites_hgraph_rmat, scale, edge_factor, seed, a, b, c, wmin, wmax,
                         n_limit, exact_edges, threads)

    @classmethod
    def rmat_batch(cls, count, seed, scale_lo=10, scale_hi=16, ef_lo=8, ef_hi=32, wmin=1,
                   wmax=255, threads=0):
        """Disjoint union of `count` seeded R-MAT graphs of uneven sizes.
        Returns (graph, vertex offsets [count+1], per-graph sources [count])."""
        voff = np.zeros(count + 1, np.uint32)
        src = np.zeros(count, np.uint32)
        out = ctypes.c_void_p()
        check(lib().kites_hgraph_rmat_batch(count, seed, scale_lo, scale_hi, ef_lo, ef_hi, wmin,
                                            wmax, threads, ctypes.byref(out),
                                            voff.ctypes.data_as(_lib.u32p),
                                            src.ctypes.data_as(_lib.u32p)))
        return HostGraph(out), voff, src

    @classmethod
    def erdos_renyi(cls, n, draws, seed, wmin=1, wmax=255, threads=0):
        return cls._make(lib().kites_hgraph_er, n, draws, seed, wmin, wmax, threads)

    @classmethod
    def grid(cls, rows, cols, seed, wmin=1, wmax=100, threads=0):
        return cls._make(lib().kites_hgraph_grid, rows, cols, seed, wmin, wmax, threads)

    @classmethod
    def from_csr(cls, R, C, E, threads=0):
        R = np.ascontiguousarray(R, np.uint32)
        C = np.ascontiguousarray(C, np.uint32)
        E = np.ascontiguousarray(E, np.uint32)
        n = len(R) - 1
        p = lambda a: a.ctypes.data_as(_lib.u32p)
        return cls._make(lib().kites_hgraph_from_csr, n, p(R), p(C), p(E), threads)

    @classmethod
    def from_edges(cls, n, src, dst, w, dedup=False, threads=0):
        """Host build from the edge list src[k] -> dst[k] with weight w[k]
        (the arrays DeviceGraph.from_edges builds on the GPU)."""
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = np.ascontiguousarray(w, np.uint32)
        if not (len(src) == len(dst) == len(w)):
            raise ValueError("src, dst and w must have one entry per edge")
        p = lambda a: a.ctypes.data_as(_lib.u32p)
        return cls._make(lib().kites_hgraph_from_edges, n, len(src), p(src), p(dst), p(w),
                         int(dedup), threads)

    def save_text(self, path):
        """The reference's prepared-graph text dump (graph_output.cuh)."""
        check(lib().kites_hgraph_save_text(self._h, str(path).encode()))

    @classmethod
    def load_text(cls, path, threads=0):
        """Reads a prepared-graph text dump (graph_csr.h:112-166)."""
        return cls._make(lib().kites_hgraph_load_text, str(path).encode(), threads)

    def save_binary(self, path):
        check(lib().kites_hgraph_save_binary(self._h, str(path).encode()))

    @classmethod
    def load_binary(cls, path):
        return cls._make(lib().kites_hgraph_load_binary, str(path).encode())

    @classmethod
    def load_edgelist(cls, path, vpg=1, s_d=True, directed=True, complete=True, source=30):
        return cls._make(lib().kites_hgraph_load_edgelist, str(path).encode(), vpg, int(s_d),
                         int(directed), int(complete), source)

    def symmetrize(self, threads=0):
        return HostGraph._make(lib().kites_hgraph_symmetrize, self._h, threads)

    def info(self):
        n, nr, m = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().kites_hgraph_info(self._h, ctypes.byref(n), ctypes.byref(nr), ctypes.byref(m)))
        return n.value, nr.value, m.value

    @property
    def n(self):
        return self.info()[0]

    @property
    def m(self):
        return self.info()[2]

    def arrays(self):
        """(R, C, E, oR, oC, oE) as numpy copies."""
        n, _, m = self.info()
        ps = [_lib.u32p() for _ in range(6)]
        check(lib().kites_hgraph_arrays(self._h, *[ctypes.byref(p) for p in ps]))
        sizes = [n + 1, m, m, n + 1, m, m]
        return tuple(_arr(p, s) for p, s in zip(ps, sizes))

    def load_arrays(self):
        """Edge-list loads: (V0, frontier bits, dir_R, dir_C)."""
        n, _, _ = self.info()
        v0, fr, dr, dc = _lib.u32p(), _lib.u8p(), _lib.u32p(), _lib.u32p()
        nb, ndr, ndc = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        check(lib().kites_hgraph_load_arrays(self._h, ctypes.byref(v0), ctypes.byref(fr),
                                             ctypes.byref(nb), ctypes.byref(dr),
                                             ctypes.byref(ndr), ctypes.byref(dc),
                                             ctypes.byref(ndc)))
        return (_arr(v0, n), _arr(fr, nb.value, np.uint8), _arr(dr, ndr.value),
                _arr(dc, ndc.value))


def run_opts(sched="auto", gran=3, forced_variant=8, alpha=14.0, use_graph=True, chunk=16,
             grid=0, max_iterations=1 << 16, keep_stats=False, margin=0.0, delta=0,
             split_pull=False, persist=True, small_grid=0, tiny_grid=0, pipeline=True):
    o = _lib.RunOpts()
    o.sched = SCHED[sched] if isinstance(sched, str) else int(sched)
    o.gran = gran
    o.forced_variant = (VARIANTS.index(forced_variant) if isinstance(forced_variant, str)
                        else int(forced_variant))
    o.alpha = alpha
    o.use_graph = int(use_graph)
    o.chunk = chunk
    o.grid = grid
    o.max_iterations = max_iterations
    o.keep_stats = int(keep_stats)
    o.margin = margin
    o.delta = delta
    o.split_pull = int(split_pull)
    o.persist = int(persist)
    o.small_grid = small_grid
    o.tiny_grid = tiny_grid
    o.pipeline = int(pipeline)
    return o


def batch_opts(edges_per_item=256, block=512, unroll=8, grid=0, use_graph=True, chunk=8,
               max_iterations=1 << 16, keep_stats=False, per_source_masks=False, top_down=1):
    o = _lib.BatchOpts()
    o.edges_per_item, o.block, o.unroll, o.grid = edges_per_item, block, unroll, grid
    o.use_graph, o.chunk = int(use_graph), chunk
    o.max_iterations, o.keep_stats = max_iterations, int(keep_stats)
    o.per_source_masks = int(per_source_masks)
    o.top_down = int(top_down)
    return o


class DeviceGraph:
    """Device-resident graph (the counterpart of make_graph_csr_for_device)."""

    def __init__(self, host, device=0):
        out = ctypes.c_void_p()
        check(lib().kites_graph_upload(device, host.handle, ctypes.byref(out)))
        self._adopt(out)

    def _adopt(self, out):
        self._g = out
        n, m = ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().kites_graph_info(self._g, ctypes.byref(n), ctypes.byref(m)))
        self.n, self.m = n.value, m.value

    @classmethod
    def from_edges(cls, n, src, dst, w, dedup=False, device=0):
        """Graph built on the GPU from the edge list src[k] -> dst[k] with
        weight w[k] (ids < n).  Returns (graph, device build time in ms)."""
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = np.ascontiguousarray(w, np.uint32)
        if not (len(src) == len(dst) == len(w)):
            raise ValueError("src, dst and w must have one entry per edge")
        out, ms = ctypes.c_void_p(), ctypes.c_float()
        check(lib().kites_graph_from_edges(device, n, len(src), src.ctypes.data_as(_lib.u32p),
                                           dst.ctypes.data_as(_lib.u32p),
                                           w.ctypes.data_as(_lib.u32p), int(dedup),
                                           ctypes.byref(out), ctypes.byref(ms)))
        g = cls.__new__(cls)
        g._adopt(out)
        return g, ms.value

    def arrays(self):
        """The device graph's arrays: (R, C, E, oR, oC, oE, nzi, nzR)."""
        n, m = self.n, self.m
        R, oR = np.empty(n + 1, np.uint32), np.empty(n + 1, np.uint32)
        C, E, oC, oE = (np.empty(m, np.uint32) for _ in range(4))
        nnz = ctypes.c_uint32()
        check(lib().kites_graph_arrays(self._g, None, None, None, None, None, None, None, None,
                                       ctypes.byref(nnz)))
        nzi, nzR = np.empty(nnz.value, np.uint32), np.empty(nnz.value + 1, np.uint32)
        p = _lib.u32p
        check(lib().kites_graph_arrays(self._g, R.ctypes.data_as(p), C.ctypes.data_as(p),
                                       E.ctypes.data_as(p), oR.ctypes.data_as(p),
                                       oC.ctypes.data_as(p), oE.ctypes.data_as(p),
                                       nzi.ctypes.data_as(p), nzR.ctypes.data_as(p),
                                       ctypes.byref(nnz)))
        return R, C, E, oR, oC, oE, nzi, nzR

    def close(self):
        if getattr(self, "_g", None) is not None and self._g.value:
            lib().kites_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, V, frontier=None):
        V = np.ascontiguousarray(V, np.uint32)
        assert len(V) == self.n
        fp = None
        if frontier is not None:
            frontier = np.ascontiguousarray(frontier, np.uint8)
            fp = frontier.ctypes.data_as(_lib.u8p)
        check(lib().kites_graph_set_state(self._g, V.ctypes.data_as(_lib.u32p), fp))

    def init(self, algo, source=0):
        check(lib().kites_graph_init(self._g, ALGO[algo], source))

    def init_sources(self, algo, sources):
        """Standard start from a set of sources (each at 0, the rest INF)."""
        src = np.ascontiguousarray(sources, np.uint32)
        check(lib().kites_graph_init_sources(self._g, ALGO[algo], src.ctypes.data_as(_lib.u32p),
                                             len(src)))

    def run(self, algo, **kw):
        o = run_opts(**kw)
        info = _lib.RunInfo()
        check(lib().kites_graph_run(self._g, ALGO[algo], ctypes.byref(o), ctypes.byref(info)))
        return dict(iterations=info.iterations, ms=info.ms, delta=bool(info.delta),
                    min_fast=bool(info.min_fast), graph_builds=info.graph_builds)

    def values(self):
        out = np.empty(self.n, np.uint32)
        check(lib().kites_graph_values(self._g, out.ctypes.data_as(_lib.u32p)))
        return out

    def stats(self):
        cnt = ctypes.c_uint32()
        check(lib().kites_graph_stats(self._g, None, 0, ctypes.byref(cnt)))
        buf = (_lib.IterStats * max(1, cnt.value))()
        check(lib().kites_graph_stats(self._g, buf, cnt.value, ctypes.byref(cnt)))
        keys = [f[0] for f in _lib.IterStats._fields_]
        return [{k: getattr(buf[i], k) for k in keys} for i in range(cnt.value)]

    def record(self, algo, rollout=0, **kw):
        """Runs `algo` from the current state (set by init / set_state) with
        a delta schedule, timing all 8 delta variants at every iteration
        (alone, or with the rest of the run under the forced `rollout`
        variant).  Returns (features [S, 9], ms [S, 8], iteration [S], own
        choice [S])."""
        kw.setdefault("sched", "plain")
        o = run_opts(**kw)
        cnt = ctypes.c_uint32()
        cap = 1 << 14
        buf = (_lib.VariantSample * cap)()
        rv = VARIANTS.index(rollout) if isinstance(rollout, str) else int(rollout)
        check(lib().kites_graph_record(self._g, ALGO[algo], ctypes.byref(o), rv, buf, cap,
                                       ctypes.byref(cnt)))
        k = min(cap, cnt.value)
        feat = np.array([list(buf[i].feat) for i in range(k)], np.float32).reshape(k, _lib.MLP_IN)
        ms = np.array([list(buf[i].ms) for i in range(k)], np.float32).reshape(k, 8)
        it = np.array([buf[i].iter for i in range(k)], np.uint32)
        pv = np.array([buf[i].plain_variant for i in range(k)], np.uint32)
        return feat, ms, it, pv

    def batch(self, algo, sources, **kw):
        """Multi-source SSSP / BFS (<= 64 sources in one run).  Returns
        (values [n, K] with column k the run from sources[k], run info)."""
        src = np.ascontiguousarray(sources, np.uint32)
        K = len(src)
        o = batch_opts(**kw)
        info = _lib.BatchInfo()
        check(lib().kites_batch_run(self._g, ALGO[algo], src.ctypes.data_as(_lib.u32p), K,
                                    ctypes.byref(o), ctypes.byref(info)))
        out = np.empty((self.n, K), np.uint32)
        if K:
            check(lib().kites_batch_values(self._g, out.ctypes.data_as(_lib.u32p)))
        return out, dict(iterations=info.iterations, ms=info.ms, graph_builds=info.graph_builds)

    def batch_run(self, algo, sources, **kw):
        """As batch() without copying the values back (timing loops)."""
        src = np.ascontiguousarray(sources, np.uint32)
        o = batch_opts(**kw)
        info = _lib.BatchInfo()
        check(lib().kites_batch_run(self._g, ALGO[algo], src.ctypes.data_as(_lib.u32p), len(src),
                                    ctypes.byref(o), ctypes.byref(info)))
        return dict(iterations=info.iterations, ms=info.ms, graph_builds=info.graph_builds)

    def autotune(self, algo, sources=(0,), **kw):
        """Times every (workgroups per CU, chunk) shape of `algo` from the
        standard start (one source, a seed list, or none for CC); returns
        (best run options as a dict, best ms, table of (per CU, block,
        chunk, ms))."""
        src = np.ascontiguousarray(sources, np.uint32)
        o = run_opts(**kw)
        ms = ctypes.c_float()
        tab = (_lib.TuneEntry * 64)()
        cnt = ctypes.c_uint32()
        check(lib().kites_graph_autotune(self._g, ALGO[algo], src.ctypes.data_as(_lib.u32p),
                                         len(src), ctypes.byref(o), ctypes.byref(ms), tab, 64,
                                         ctypes.byref(cnt)))
        best = {"grid": o.grid, "chunk": o.chunk}
        table = [(t.edges_per_item, t.block, t.unroll, t.ms) for t in tab[:cnt.value]]
        return best, ms.value, table

    def bfs256(self, sources, copy=True, **kw):
        """Bit-parallel BFS from up to 256 sources.  Returns (levels [n, K]
        or None when copy is False, run info)."""
        src = np.ascontiguousarray(sources, np.uint32)
        K = len(src)
        o = batch_opts(**kw)
        info = _lib.BatchInfo()
        check(lib().kites_bfs256_run(self._g, src.ctypes.data_as(_lib.u32p), K, ctypes.byref(o),
                                     ctypes.byref(info)))
        out = None
        if copy:
            out = np.empty((self.n, K), np.uint32)
            if K:
                check(lib().kites_bfs256_levels(self._g, out.ctypes.data_as(_lib.u32p)))
        return out, dict(iterations=info.iterations, ms=info.ms, graph_builds=info.graph_builds)

    def bfs256_autotune(self, sources, repeats=2, **kw):
        """Times every (edges_per_item, block) shape of the bit-parallel BFS;
        returns (best options dict, best ms, table of (edges_per_item, block,
        ms)).  Every shape must give identical levels."""
        src = np.ascontiguousarray(sources, np.uint32)
        o = batch_opts(**kw)
        ms = ctypes.c_float()
        tab = (_lib.TuneEntry * 64)()
        cnt = ctypes.c_uint32()
        check(lib().kites_bfs256_autotune(self._g, src.ctypes.data_as(_lib.u32p), len(src),
                                          repeats, ctypes.byref(o), ctypes.byref(ms), tab, 64,
                                          ctypes.byref(cnt)))
        best = {f[0]: getattr(o, f[0]) for f in _lib.BatchOpts._fields_}
        table = [(t.edges_per_item, t.block, t.ms) for t in tab[:cnt.value]]
        return best, ms.value, table

    def bfs256_stats(self):
        cnt = ctypes.c_uint32()
        check(lib().kites_bfs256_stats(self._g, None, 0, ctypes.byref(cnt)))
        buf = (_lib.Bfs256IterStats * max(1, cnt.value))()
        check(lib().kites_bfs256_stats(self._g, buf, cnt.value, ctypes.byref(cnt)))
        keys = ["updated", "push", "active_edges", "new_pairs", "out_edges", "max_out"]
        return [{k: getattr(buf[i], k) for k in keys} for i in range(cnt.value)]

    def batch_stats(self):
        cnt = ctypes.c_uint32()
        check(lib().kites_batch_stats(self._g, None, 0, ctypes.byref(cnt)))
        buf = (_lib.BatchIterStats * max(1, cnt.value))()
        check(lib().kites_batch_stats(self._g, buf, cnt.value, ctypes.byref(cnt)))
        keys = [f[0] for f in _lib.BatchIterStats._fields_]
        return [{k: getattr(buf[i], k) for k in keys} for i in range(cnt.value)]

    def batch_autotune(self, algo, sources, repeats=2, **kw):
        """Times every work-item shape; returns (best options dict, best ms,
        table of (edges_per_item, block, unroll, ms))."""
        src = np.ascontiguousarray(sources, np.uint32)
        o = batch_opts(**kw)
        ms = ctypes.c_float()
        tab = (_lib.TuneEntry * 64)()
        cnt = ctypes.c_uint32()
        check(lib().kites_batch_autotune(self._g, ALGO[algo], src.ctypes.data_as(_lib.u32p),
                                         len(src), repeats, ctypes.byref(o), ctypes.byref(ms),
                                         tab, 64, ctypes.byref(cnt)))
        best = {f[0]: getattr(o, f[0]) for f in _lib.BatchOpts._fields_}
        table = [(t.edges_per_item, t.block, t.unroll, t.ms) for t in tab[:cnt.value]]
        return best, ms.value, table

    def set_model(self, weights):
        w = np.ascontiguousarray(weights, np.float32)
        check(lib().kites_graph_set_model(self._g, w.ctypes.data_as(_lib.f32p), len(w)))

    def pagerank(self, damping=0.85, tol=1e-4, max_iterations=0):
        """Float PageRank in the ordered mode (defined bit for bit).
        Returns (values as float32, run info)."""
        out = np.empty(self.n, np.float32)
        info = _lib.RunInfo()
        check(lib().kites_pagerank_f32(self._g, damping, tol, max_iterations,
                                       out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       ctypes.byref(info)))
        return out, dict(iterations=info.iterations, ms=info.ms)

    def sssp_noalloc(self, source, **kw):
        """kites_sssp_u32_noalloc: SSSP with the run state in the library's
        fixed device array (no allocation).  Returns (distances, run info)."""
        o = run_opts(**kw)
        info = _lib.RunInfo()
        out = np.empty(self.n, np.uint32)
        check(lib().kites_sssp_u32_noalloc(self._g, source, ctypes.byref(o), ctypes.byref(info),
                                           out.ctypes.data_as(_lib.u32p)))
        return out, dict(iterations=info.iterations, ms=info.ms, delta=bool(info.delta),
                         min_fast=bool(info.min_fast))

    def solve(self, algo, source=0, **kw):
        """Standard init + run; returns (values, run info)."""
        self.init(algo, source)
        info = self.run(algo, **kw)
        return self.values(), info


def autotune_vpg(path, source=30, vpgs=(1, 2, 4, 8, 16, 32, 64), algo="sssp_wrap", repeats=3,
                 device=0, **kw):
    """Vertex-per-group of the reference algorithm (the classic schedule) for
    an edge-list graph: the list is loaded at every vpg in `vpgs` (the
    reference builds the group out-lists there, kites/graph/graph_raw.h:
    142-151, and picks its kernel by vpg, kites/device/compute.cuh:93-117),
    run `repeats` times from the loader's initial values and frontier with
    the reference example's callback for `source`, and the fastest device
    time is kept.  Every vpg must give the same values (RuntimeError
    otherwise).  Returns (best vpg, its ms, {vpg: ms})."""
    table, ref = {}, None
    for vpg in vpgs:
        h = HostGraph.load_edgelist(path, vpg=vpg, source=source)
        V0, F0 = h.load_arrays()[:2]
        d = DeviceGraph(h, device)
        best = None
        for _ in range(max(1, repeats)):
            d.set_state(V0, F0)
            info = d.run(algo, sched="classic", **kw)
            best = info["ms"] if best is None else min(best, info["ms"])
        v = d.values()
        d.close()
        if ref is None:
            ref = v
        elif not np.array_equal(v, ref):
            raise RuntimeError(f"vertex-per-group {vpg} gives different values than {vpgs[0]}")
        table[vpg] = best
    b = min(table, key=table.get)
    return b, table[b], table


def mlp_nparam():
    return int(lib().kites_mlp_nparam())


def device_count():
    c = ctypes.c_int32()
    check(lib().kites_device_count(ctypes.byref(c)))
    return c.value
