"""The C ABI library without a GPU: it loads, exports every symbol declared
in include/ogsx.h, and its host-side CRS builder (no device work) equals the
oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from ogs5_amd import mesh as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "ogsx.h")).read()
    return sorted(set(re.findall(r"\b(ogsx_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol(ox):
    from ogs5_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES)


def test_version_and_device_count(ox):
    assert ox.lib().ogsx_version().decode().startswith("ogsx")
    assert ox.device_count() >= 0


@pytest.mark.parametrize("mk", [lambda: M.quad_mesh(13, 9), lambda: M.hex_mesh(5, 4, 3),
                                lambda: M.tet_mesh(4, 3, 5), lambda: M.line_mesh(7)])
def test_host_pattern_bit_exact_vs_oracle(ox, ffi, mk):
    m = mk()
    p = ox.Pattern(m.n_nodes, m.etype, m.conn)
    got = p.export()
    exp = ffi.pattern(m.n_nodes, m.conn)
    for g, e in zip(got, exp):
        np.testing.assert_array_equal(g, e)
    # colouring: no two elements of one colour share a node
    col = p.colors()
    for c in range(p.n_colors):
        nodes = m.conn[col == c].ravel()
        assert np.unique(nodes).size == nodes.size


def test_pattern_rejects_bad_input(ox):
    with pytest.raises(ox.OgsxError):
        ox.Pattern(4, ox.QUAD4, np.array([[0, 1, 2, 7]]))
    with pytest.raises(ox.OgsxError):
        ox.Pattern(4, 9, np.array([[0, 1, 2, 3]]))


def test_system_create_without_gpu_fails_loudly(ox):
    if ox.device_count() > 0:
        pytest.skip("a GPU is present")
    m = M.quad_mesh(2, 2)
    p = ox.Pattern(m.n_nodes, m.etype, m.conn)
    with pytest.raises(ox.OgsxError, match="no HIP device"):
        ox.System(p)


def test_system_create_rejects_empty_rows(ox):
    """A node in no element has an empty CRS row (the oracle's pattern has
    the same row and no diagonal, diag = -1): refused before any device work,
    so it needs no GPU."""
    from oracle import ffi
    conn = np.array([[0, 1], [1, 2]], np.int64)
    p = ox.Pattern(5, ox.LINE2, conn)
    ptr, col, diag = p.export()
    assert list(diag) == list(ffi.pattern(5, conn)[2]) == [0, 3, 6, -1, -1]
    with pytest.raises(ox.OgsxError, match="node 3 belongs to no element"):
        ox.System(p)
