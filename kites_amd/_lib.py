"""ctypes binding of ``libkites_amd.so`` (the C ABI in ``include/kites_amd.h``).

Loading fails loudly when the library has not been built: there is no
Python or CPU fallback for any engine entry point.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# KITES_AMD_LIB names another build of the same library (A/B of build options).
LIB_PATH = os.environ.get("KITES_AMD_LIB") or os.path.join(HERE, "libkites_amd.so")

KITES_OK = 0
KITES_ERR_ARG = 1
ALGO = {"sssp": 0, "bfs": 1, "cc": 2, "sssp_wrap": 3, "widest": 4}
SCHED = {"auto": 0, "classic": 1, "plain": 2, "forced": 3, "learned": 4}
VARIANTS = ["classic", "push_t", "push_w", "push_b", "push_lb", "pull_t", "pull_w", "pull_b",
            "pull_lb", "init_lb", "split", "persist"]
INF = 0xFFFFFFFF
MLP_IN = 10   # scheduler features (kites_variant_sample.feat)


class KitesError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kites error {code}: {msg}")
        self.code = code


class RunOpts(ctypes.Structure):
    _fields_ = [("sched", ctypes.c_int32), ("gran", ctypes.c_int32),
                ("forced_variant", ctypes.c_int32), ("alpha", ctypes.c_float),
                ("use_graph", ctypes.c_int32), ("chunk", ctypes.c_int32),
                ("grid", ctypes.c_int32), ("max_iterations", ctypes.c_uint32),
                ("keep_stats", ctypes.c_int32), ("margin", ctypes.c_float),
                ("delta", ctypes.c_uint32), ("split_pull", ctypes.c_int32),
                ("persist", ctypes.c_int32), ("small_grid", ctypes.c_int32),
                ("tiny_grid", ctypes.c_int32), ("pipeline", ctypes.c_int32)]


class RunInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint32), ("ms", ctypes.c_float),
                ("delta", ctypes.c_int32), ("min_fast", ctypes.c_int32),
                ("graph_builds", ctypes.c_uint32)]


class IterStats(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint32), ("variant", ctypes.c_uint32),
                ("max_outdeg", ctypes.c_uint32), ("in_count", ctypes.c_uint32),
                ("sum_outdeg", ctypes.c_uint64), ("work_edges", ctypes.c_uint64),
                ("gathers", ctypes.c_uint64), ("active", ctypes.c_uint64),
                ("qe", ctypes.c_uint64)]


class BatchOpts(ctypes.Structure):
    _fields_ = [("edges_per_item", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("unroll", ctypes.c_uint32), ("grid", ctypes.c_int32),
                ("use_graph", ctypes.c_int32), ("chunk", ctypes.c_int32),
                ("max_iterations", ctypes.c_uint32), ("keep_stats", ctypes.c_int32),
                ("per_source_masks", ctypes.c_int32), ("top_down", ctypes.c_int32)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_uint32), ("ms", ctypes.c_float),
                ("graph_builds", ctypes.c_uint32)]


class BatchIterStats(ctypes.Structure):
    _fields_ = [("updated", ctypes.c_uint32), ("flushes", ctypes.c_uint32),
                ("active_edges", ctypes.c_uint64), ("gathered", ctypes.c_uint64),
                ("moved", ctypes.c_uint64)]


class Bfs256IterStats(ctypes.Structure):
    _fields_ = [("updated", ctypes.c_uint32), ("push", ctypes.c_uint32),
                ("active_edges", ctypes.c_uint64), ("new_pairs", ctypes.c_uint64),
                ("out_edges", ctypes.c_uint64), ("max_out", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class TuneEntry(ctypes.Structure):
    _fields_ = [("edges_per_item", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("unroll", ctypes.c_uint32), ("ms", ctypes.c_float)]


class VariantSample(ctypes.Structure):
    _fields_ = [("feat", ctypes.c_float * MLP_IN), ("ms", ctypes.c_float * 8),
                ("iter", ctypes.c_uint32), ("plain_variant", ctypes.c_uint32)]


_lib = None

u32p = ctypes.POINTER(ctypes.c_uint32)
u8p = ctypes.POINTER(ctypes.c_uint8)
f32p = ctypes.POINTER(ctypes.c_float)
vp = ctypes.c_void_p
vpp = ctypes.POINTER(ctypes.c_void_p)


def _sig(L, name, args, res=ctypes.c_int):
    f = getattr(L, name)
    f.argtypes = args
    f.restype = res


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is not built; run `python -m kites_amd.build` (no fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    _sig(L, "kites_last_error", [], ctypes.c_char_p)
    _sig(L, "kites_device_count", [ctypes.POINTER(ctypes.c_int32)])
    _sig(L, "kites_run_opts_default", [ctypes.POINTER(RunOpts)], None)
    _sig(L, "kites_hgraph_rmat", [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.c_uint64, ctypes.c_int32, vpp])
    _sig(L, "kites_hgraph_rmat_batch", [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, vpp,
                                        u32p, u32p])
    _sig(L, "kites_graph_init_sources", [vp, ctypes.c_int32, u32p, ctypes.c_uint32])
    _sig(L, "kites_hgraph_er", [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint64,
                                ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, vpp])
    _sig(L, "kites_hgraph_grid", [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, vpp])
    _sig(L, "kites_hgraph_symmetrize", [vp, ctypes.c_int32, vpp])
    _sig(L, "kites_hgraph_from_csr", [ctypes.c_uint32, u32p, u32p, u32p, ctypes.c_int32, vpp])
    _sig(L, "kites_hgraph_load_edgelist", [ctypes.c_char_p, ctypes.c_uint32, ctypes.c_int32,
                                           ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, vpp])
    _sig(L, "kites_hgraph_info", [vp, u32p, u32p, ctypes.POINTER(ctypes.c_uint64)])
    _sig(L, "kites_hgraph_arrays", [vp] + [ctypes.POINTER(u32p)] * 6)
    _sig(L, "kites_hgraph_load_arrays", [vp, ctypes.POINTER(u32p), ctypes.POINTER(u8p),
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(u32p),
                                         ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(u32p),
                                         ctypes.POINTER(ctypes.c_uint64)])
    _sig(L, "kites_hgraph_destroy", [vp], None)
    _sig(L, "kites_graph_upload", [ctypes.c_int32, vp, vpp])
    _sig(L, "kites_graph_destroy", [vp], None)
    _sig(L, "kites_graph_info", [vp, u32p, ctypes.POINTER(ctypes.c_uint64)])
    _sig(L, "kites_graph_from_edges", [ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint64, u32p,
                                       u32p, u32p, ctypes.c_int32, vpp,
                                       ctypes.POINTER(ctypes.c_float)])
    _sig(L, "kites_graph_arrays", [vp] + [u32p] * 9)
    _sig(L, "kites_pagerank_f32", [vp, ctypes.c_float, ctypes.c_float, ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_float), ctypes.POINTER(RunInfo)])
    _sig(L, "kites_hgraph_save_text", [vp, ctypes.c_char_p])
    _sig(L, "kites_hgraph_load_text", [ctypes.c_char_p, ctypes.c_int32, vpp])
    _sig(L, "kites_hgraph_save_binary", [vp, ctypes.c_char_p])
    _sig(L, "kites_hgraph_load_binary", [ctypes.c_char_p, vpp])
    _sig(L, "kites_hgraph_from_edges", [ctypes.c_uint32, ctypes.c_uint64, u32p, u32p, u32p,
                                        ctypes.c_int32, ctypes.c_int32, vpp])
    _sig(L, "kites_graph_set_state", [vp, u32p, u8p])
    _sig(L, "kites_graph_init", [vp, ctypes.c_int32, ctypes.c_uint32])
    _sig(L, "kites_graph_run", [vp, ctypes.c_int32, ctypes.POINTER(RunOpts),
                                ctypes.POINTER(RunInfo)])
    _sig(L, "kites_graph_values", [vp, u32p])
    _sig(L, "kites_graph_stats", [vp, ctypes.POINTER(IterStats), ctypes.c_uint32, u32p])
    for nm in ("kites_sssp_u32", "kites_bfs_u32", "kites_sssp_u32_noalloc"):
        _sig(L, nm, [vp, ctypes.c_uint32, ctypes.POINTER(RunOpts), ctypes.POINTER(RunInfo), u32p])
    _sig(L, "kites_cc_u32", [vp, ctypes.POINTER(RunOpts), ctypes.POINTER(RunInfo), u32p])
    _sig(L, "kites_mlp_nparam", [], ctypes.c_uint32)
    _sig(L, "kites_graph_set_model", [vp, f32p, ctypes.c_uint32])
    _sig(L, "kites_graph_record", [vp, ctypes.c_int32, ctypes.POINTER(RunOpts), ctypes.c_uint32,
                                   ctypes.POINTER(VariantSample), ctypes.c_uint32, u32p])
    _sig(L, "kites_batch_opts_default", [ctypes.POINTER(BatchOpts)], None)
    _sig(L, "kites_batch_run", [vp, ctypes.c_int32, u32p, ctypes.c_uint32,
                                ctypes.POINTER(BatchOpts), ctypes.POINTER(BatchInfo)])
    _sig(L, "kites_batch_values", [vp, u32p])
    _sig(L, "kites_batch_stats", [vp, ctypes.POINTER(BatchIterStats), ctypes.c_uint32, u32p])
    _sig(L, "kites_batch_autotune", [vp, ctypes.c_int32, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.POINTER(BatchOpts), f32p,
                                     ctypes.POINTER(TuneEntry), ctypes.c_uint32, u32p])
    _sig(L, "kites_bfs256_run", [vp, u32p, ctypes.c_uint32, ctypes.POINTER(BatchOpts),
                                 ctypes.POINTER(BatchInfo)])
    _sig(L, "kites_bfs256_levels", [vp, u32p])
    _sig(L, "kites_bfs256_stats", [vp, ctypes.POINTER(Bfs256IterStats), ctypes.c_uint32, u32p])
    _sig(L, "kites_bfs256_autotune", [vp, u32p, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.POINTER(BatchOpts), f32p, ctypes.POINTER(TuneEntry),
                                      ctypes.c_uint32, u32p])
    _sig(L, "kites_graph_autotune", [vp, ctypes.c_int32, u32p, ctypes.c_uint32,
                                     ctypes.POINTER(RunOpts), f32p, ctypes.POINTER(TuneEntry),
                                     ctypes.c_uint32, u32p])
    _sig(L, "kites_mlp_last_error", [], ctypes.c_char_p)
    _sig(L, "kites_mlp_forward", [ctypes.c_int32, f32p, f32p, ctypes.c_uint32, f32p])
    _sig(L, "kites_mlp_loss_grad", [ctypes.c_int32, f32p, f32p, f32p, f32p, ctypes.c_uint32,
                                    f32p, f32p])
    _sig(L, "kites_mlp_train", [ctypes.c_int32, f32p, f32p, f32p, f32p, ctypes.c_uint32,
                                ctypes.c_uint32, ctypes.c_float, f32p, f32p])
    _lib = L
    return L


def check(rc):
    if rc != KITES_OK:
        raise KitesError(rc, lib().kites_last_error().decode(errors="replace"))


def check_mlp(rc):
    if rc != KITES_OK:
        raise KitesError(rc, lib().kites_mlp_last_error().decode(errors="replace"))


# Every C-ABI symbol include/kites_amd.h declares (checked by the CPU tests).
EXPORTED = [
    "kites_last_error", "kites_device_count", "kites_run_opts_default",
    "kites_hgraph_rmat", "kites_hgraph_er", "kites_hgraph_grid", "kites_hgraph_symmetrize",
    "kites_hgraph_from_csr", "kites_hgraph_load_edgelist", "kites_hgraph_info",
    "kites_hgraph_arrays", "kites_hgraph_load_arrays", "kites_hgraph_destroy",
    "kites_graph_upload", "kites_graph_destroy", "kites_graph_info", "kites_graph_set_state",
    "kites_graph_init", "kites_graph_run", "kites_graph_values", "kites_graph_stats",
    "kites_sssp_u32", "kites_bfs_u32", "kites_cc_u32", "kites_mlp_nparam",
    "kites_graph_set_model", "kites_graph_record", "kites_mlp_last_error", "kites_mlp_forward",
    "kites_mlp_loss_grad", "kites_mlp_train", "kites_batch_opts_default", "kites_batch_run",
    "kites_batch_values", "kites_batch_stats", "kites_batch_autotune", "kites_bfs256_run",
    "kites_bfs256_levels", "kites_bfs256_stats", "kites_bfs256_autotune", "kites_hgraph_rmat_batch",
    "kites_graph_init_sources", "kites_graph_autotune", "kites_graph_from_edges",
    "kites_graph_arrays", "kites_hgraph_from_edges", "kites_hgraph_save_text",
    "kites_hgraph_load_text", "kites_hgraph_save_binary", "kites_hgraph_load_binary",
    "kites_pagerank_f32", "kites_sssp_u32_noalloc",
]
