"""Writes the committed fixtures of this directory.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_golden.py

* wikivote.txt.gz -- the reference's only data file,
  examples/Sample-Graph-Wiki-Vote.txt (SNAP Wiki-Vote, CRLF, 4 '#' lines),
  byte for byte, gzip-compressed (the reference is not on the GPU box).
* wikivote_sssp_src30.npz -- the reference example's SSSP (source 30,
  INF = 1073741823, weights (s+1)*(d+2)%127, functor srcV + w) after graph
  construction as graph_raw does it: `V` (8,320 u32).  Written only if the
  oracle's Gauss-Seidel fixpoint and an independent Dijkstra agree on it.
* c1_digest.json -- the C1 graph (R-MAT, 7,115 vertices, 103,689 edges,
  weights U[1,64], seed 1) counts, sha256 of its in-CSR and of its SSSP
  result from the max out-degree vertex.
"""
import gzip
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import graph_raw_oracle as gro  # noqa: E402
from oracle import liboracle as O  # noqa: E402

REF_FILE = "/root/reference/examples/Sample-Graph-Wiki-Vote.txt"
C1 = dict(scale=13, seed=1, wmin=1, wmax=64, n_limit=7115, exact_edges=103689)


def wikivote():
    raw = open(REF_FILE, "rb").read()
    with gzip.open(os.path.join(HERE, "wikivote.txt.gz"), "wb", compresslevel=9) as f:
        f.write(raw)
    g = gro.build(raw.decode(), gro.example_callback(30), 1, True, True, True)
    V, sweeps = O.fixpoint(g["R"], g["C"], g["E"], g["V"], O.SSSP_WRAP)
    d = O.dijkstra(g["R"], g["C"], g["E"], 30)
    exp = g["V"].copy()
    reach = d != O.INF
    exp[reach] = np.minimum(exp[reach], d[reach])
    assert np.array_equal(V, exp), "oracle and Dijkstra disagree"
    np.savez_compressed(os.path.join(HERE, "wikivote_sssp_src30.npz"), V=V)
    print("wikivote: n", g["n"], "m", g["m"], "sweeps", sweeps, "reached", int(reach.sum()))


def c1():
    import kites_amd as K
    h = K.HostGraph.rmat(C1["scale"], 0, C1["seed"], wmin=C1["wmin"], wmax=C1["wmax"],
                         n_limit=C1["n_limit"], exact_edges=C1["exact_edges"])
    R, C, E, oR, oC, oE = h.arrays()
    n = len(R) - 1
    src = int(np.argmax(np.diff(oR)))
    V0, _ = O.sssp_init(n, src, R, C)
    V, _ = O.fixpoint(R, C, E, V0, O.SSSP_SAT)
    assert np.array_equal(V, O.dijkstra(R, C, E, src))
    sha = lambda *a: hashlib.sha256(b"".join(np.ascontiguousarray(x).tobytes() for x in a)).hexdigest()
    out = dict(params=C1, n=n, m=int(len(C)), source=src, csr_sha256=sha(R, C, E),
               sssp_sha256=sha(V), reached=int((V != O.INF).sum()))
    json.dump(out, open(os.path.join(HERE, "c1_digest.json"), "w"), indent=1)
    print("c1:", out)


if __name__ == "__main__":
    wikivote()
    c1()
