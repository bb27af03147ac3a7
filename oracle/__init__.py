"""CPU oracle for the vertex-centric iteration engine.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s ``cpu_baseline`` leg may import this package, and only as the
checker (or, for ``cpu_baseline``, as the timed CPU baseline).  The product
(``kites_amd``) never imports it.

Parts and what pins them:

* ``liboracle`` (``kites_oracle.c``): the reference CPU path restated
  (Gauss-Seidel fixpoint, ``kites/graph/sanity_check.cuh:21-47``) plus
  textbook Dijkstra / BFS / union-find and a multi-core frontier baseline.
  Pinned in ``tests/test_oracle.py`` by textbook algorithms, exhaustive
  Bellman-Ford on tiny graphs, closed forms and the Wiki-Vote golden vector.
* ``graph_raw_oracle``: the reference edge-list parser and CSR builder
  restated (``kites/graph/graph_raw.h:58-190``, ``graph_csr.h:45-110``).
  Pinned in ``tests/test_graph_build.py`` against ``oracle/_ref`` (the
  reference's own ``graph_raw.h`` compiled here) where that is built, and
  against committed counts of the Wiki-Vote fixture everywhere.
* ``mlp_oracle``: float64 forward/backward of the 2-layer scheduler MLP.
  Pinned by central finite differences in ``tests/test_mlp_oracle.py``.
"""
