import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: larger CPU case")


def case(msg):
    """Prints and flushes what the next GPU step runs, so a failure that
    takes the process down still names its case."""
    print(f"[case] {msg}", flush=True)


TINY_GRAPHS = ("edgeless n=70", "self-loop only", "loops and parallel edges n=70",
               "saturating weights n=5")


def tiny_graph(name):
    """Small graphs with the shapes that break index arithmetic: no edges at
    all (70 vertices, not a multiple of the 64-vertex bitmap word), a
    self-loop as the only edge, a 70-vertex graph with self-loops,
    parallel edges of different weights, a cycle of zero-weight edges, a
    vertex with in-edges only and an unreachable tail, and a 5-vertex graph
    whose path sums run past 2^32 - 1 (saturating at the unreached value;
    one edge weighs 2^32 - 1 itself)."""
    import numpy as np
    import kites_amd as K
    if name == "edgeless n=70":
        return K.HostGraph.from_csr(np.zeros(71, np.uint32), np.zeros(0, np.uint32),
                                    np.zeros(0, np.uint32))
    if name == "self-loop only":
        return K.HostGraph.from_csr(np.array([0, 1], np.uint32), np.array([0], np.uint32),
                                    np.array([3], np.uint32))
    if name == "saturating weights n=5":
        n = 5
        src = [0, 1, 0, 2, 3, 0]
        dst = [1, 2, 2, 3, 4, 4]
        w = [0xFFFFFFF0, 0x20, 0xFFFFFFFE, 5, 1, 0xFFFFFFFF]
    else:
        assert name == "loops and parallel edges n=70", name
        n = 70
        src = [0, 0, 0, 1, 1, 2, 3, 3, 5, 5, 5, 8, 9, 40, 41, 42, 63, 64, 64]
        dst = [0, 1, 1, 2, 1, 3, 4, 63, 5, 6, 6, 9, 8, 41, 42, 40, 64, 65, 69]
        w = [1, 7, 2, 4, 1, 9, 3, 250, 2, 8, 3, 1, 1, 0, 0, 0, 6, 0, 1]
    order = np.lexsort((src, dst))
    s = np.array(src, np.uint32)[order]
    t = np.array(dst, np.int64)[order]
    R = np.zeros(n + 1, np.int64)
    np.add.at(R, t + 1, 1)
    return K.HostGraph.from_csr(np.cumsum(R).astype(np.uint32), s, np.array(w, np.uint32)[order])
