"""The oracle's Gauss tables, shape functions and dense Matrix::multi against
the reference's own FEM/mathlib.cpp and FEM/matrix_class.cpp.

tests/golden/ref_mathlib_dump.txt is the output of oracle/_ref/ogs5_ref_dump,
a driver (oracle/ref_dump.cpp) linked against those two reference files
compiled from the reference tree by oracle/Makefile.  When the reference tree
is present the dump is regenerated and must equal the committed fixture.
"""
import ctypes as C
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_mathlib_dump.txt")


def _hex(s):
    return float.fromhex(s)


def _lines():
    return open(GOLDEN).read().splitlines()


def test_fixture_matches_live_reference_build(ffi):
    live = ffi.ref_dump_lines()
    if live is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    assert live == _lines()


def test_gauss_tables_bit_exact(ffi):
    L = ffi.lib()
    seen = 0
    for ln in _lines():
        t = ln.split()
        if t[0] != "gauss":
            continue
        grd, pkt = int(t[1]), int(t[2])
        assert L.orc_gauss_pkt(grd, pkt) == _hex(t[3])
        assert L.orc_gauss_fkt(grd, pkt) == _hex(t[4])
        seen += 1
    assert seen == 10


def test_tet5_rule_bit_exact(ffi):
    L = ffi.lib()
    sp = np.zeros(4)
    for ln in _lines():
        t = ln.split()
        if t[0] != "tet5":
            continue
        L.orc_sample_tet5(int(t[1]), sp.ctypes.data_as(C.c_void_p))
        assert list(sp) == [_hex(v) for v in t[2:6]]


@pytest.mark.parametrize("etype,npe,edim", [(1, 2, 1), (2, 4, 2), (3, 8, 3), (5, 4, 3)])
def test_shape_pool_bit_exact(ffi, etype, npe, edim):
    L = ffi.lib()
    rows = {}
    for ln in _lines():
        t = ln.split()
        if t[0] == "shape" and int(t[1]) == etype:
            rows.setdefault(int(t[2]), {})[int(t[3])] = [_hex(v) for v in t[4:]]
    assert rows
    for ng, pts in rows.items():
        ngp = L.orc_ngp(etype, ng)
        assert ngp == len(pts)
        N = np.zeros(ngp * npe)
        dN = np.zeros(ngp * npe * edim)
        L.orc_shape_pool(etype, ng, N.ctypes.data_as(C.c_void_p), dN.ctypes.data_as(C.c_void_p))
        for gp, vals in pts.items():
            assert list(N[gp * npe:(gp + 1) * npe]) == vals[:npe]
            assert list(dN[gp * npe * edim:(gp + 1) * npe * edim]) == vals[npe:]


def test_dense_multi_matches_reference(ffi):
    """Matrix::multi(vec, vec_result, fac) layout and accumulation."""
    L = ffi.lib()
    for ln in _lines():
        t = ln.split()
        if t[0] != "multi":
            continue
        n = int(t[1])
        c = {3: 0, 4: 1, 8: 2}[n]
        m = np.array([[1.0 / (1.0 + i + 2.0 * j + c) - 0.3 * (i == j) for j in range(n)]
                      for i in range(n)])
        x = np.array([0.1 * (i + 1) + 1.0 / (3.0 + i + c) for i in range(n)])
        b = np.array([0.25 * i for i in range(n)])
        L.orc_dense_multi(m.ctypes.data_as(C.c_void_p), n, n, x.ctypes.data_as(C.c_void_p),
                          b.ctypes.data_as(C.c_void_p), -0.7)
        assert list(b) == [_hex(v) for v in t[2:]]
