"""Lifetime and initial-state rules of the C ABI (no counterpart in the
reference, whose Linear_EQS aliases the mesh-owned SparseTable,
FEM/matrix_class.cpp:1095-1101)."""
import numpy as np
import pytest

import ogs5_amd as ox
from ogs5_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_pattern_outlives_its_systems():
    """ogsx_pattern_destroy refuses (OGSX_ESTATE) while a system created on
    the pattern exists; once the systems are gone it frees the pattern."""
    sc = W.quad_flow(6, 5, seed=1, n_steps=1)
    pat = ox.Pattern(sc.n, sc.mesh.etype, sc.mesh.conn)
    s1 = ox.System(pat, ox.FP64)
    s2 = ox.System(pat, ox.FP32_IR)
    with pytest.raises(ox.OgsxError, match="still used by 2"):
        pat.close()
    s1.close()
    with pytest.raises(ox.OgsxError, match="still used by 1"):
        pat.close()
    s2.close()
    pat.close()
    assert pat.h is None


def test_new_system_vectors_start_at_zero():
    """Every node-value vector of a new system, the stored vector included,
    starts at zero, so copying the stored vector before it was set copies
    zeros rather than stale device memory."""
    sc = W.quad_flow(6, 5, seed=1, n_steps=1)
    pat = ox.Pattern(sc.n, sc.mesh.etype, sc.mesh.conn)
    s = ox.System(pat, ox.FP64)
    for which in (ox.VEC_X, ox.VEC_B, ox.VEC_UOLD, ox.VEC_IC):
        assert not np.any(s.get_vector(which)), which
    s.set_vector(ox.VEC_X, np.ones(sc.n))
    s.copy_vector(ox.VEC_X, ox.VEC_IC)
    assert not np.any(s.get_vector(ox.VEC_X))
    s.close()
