"""The C-ABI library loads on a CPU-only host, exports every entry point
include/kites_amd.h declares, and turns bad arguments into error codes."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

K = pytest.importorskip("kites_amd")
from kites_amd import _lib  # noqa: E402


def header_functions():
    txt = open(os.path.join(ROOT, "include", "kites_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(kites_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = header_functions()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(_lib.EXPORTED) == declared


def test_mlp_param_count():
    assert K.mlp_nparam() == 10 * 32 + 32 + 8 * 32 + 8


def test_errors_become_codes():
    L = _lib.lib()
    out = ctypes.c_void_p()
    R = np.array([0, 1, 2], np.uint32)
    C = np.array([5, 0], np.uint32)          # source 5 out of range for n = 2
    E = np.array([1, 1], np.uint32)
    p = lambda a: a.ctypes.data_as(_lib.u32p)
    rc = L.kites_hgraph_from_csr(2, p(R), p(C), p(E), 0, ctypes.byref(out))
    assert rc == 1 and b"out of range" in L.kites_last_error()
    rc = L.kites_hgraph_rmat(40, 16, 1, 0.57, 0.19, 0.19, 1, 2, 0, 0, 0, ctypes.byref(out))
    assert rc != 0
    rc = L.kites_hgraph_rmat(8, 4, 1, 0.9, 0.19, 0.19, 1, 2, 0, 0, 0, ctypes.byref(out))
    assert rc == 1
    assert L.kites_graph_run(None, 0, None, None) == 1


def test_device_count_without_gpu_is_not_an_error():
    c = ctypes.c_int32(-1)
    assert _lib.lib().kites_device_count(ctypes.byref(c)) == 0
    assert c.value >= 0
