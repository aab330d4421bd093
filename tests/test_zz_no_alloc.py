"""The zero-allocation solve (ogsx_solve_crs_noalloc) against ogsx_solve.

The same FP64 recurrences on the same grid of blocks: for the same CRS
matrix, right-hand side and initial guess, both entry points must return the
same iterates bit for bit, the same iteration count, residual, return value
and bNorm.  Two calls in a row, the second on a smaller system, check that
nothing of the first call's contents of the shared device array leaks into
the second.
"""
import numpy as np
import pytest

import ogs5_amd as ox
from ogs5_amd import workloads as W

from devbuf import DevArray

pytestmark = pytest.mark.gpu


def _assembled(sc, method):
    """A system assembled for the scenario's first step (sources and
    Dirichlet applied); returns it with the CRS arrays, A, b and x0."""
    pat, s = W.make_system(sc)
    st = sc.steps[0]
    s.set_vector(ox.VEC_UOLD, sc.u0)
    s.set_vector(ox.VEC_X, sc.u0)
    s.initialize()
    s.assemble(st.dt, st.theta, st.steady)
    if sc.st_idx.size:
        s.add_rhs(sc.st_idx, sc.st_val)
    mode = ox.DIRICHLET_SYMMETRIC if method == ox.CG else ox.DIRICHLET_ROW
    s.set_dirichlet(sc.bc_idx, sc.bc_val, mode)
    ptr, col, _ = pat.export()
    return pat, s, ptr.astype(np.int32), col.astype(np.int32), s.get_matrix(), \
        s.get_vector(ox.VEC_B), s.get_vector(ox.VEC_X)


def _both(sc, method, tol=1e-12):
    pat, s, ptr, col, a, b, x0 = _assembled(sc, method)
    dev = [DevArray(v) for v in (ptr, col, a, b, x0)]
    print(f"[noalloc] {sc.name} n={sc.n} nnz={a.size} method={method}", flush=True)
    info_na, bn = ox.solve_crs_noalloc(sc.n, *[d.ptr for d in dev], bnorm=1.0, method=method,
                                       tol=tol, max_iter=20000)
    x_na = dev[4].get()
    info = s.solve(method=method, tol=tol, max_iter=20000)
    x = s.get_vector(ox.VEC_X)
    print(f"[noalloc]   system format {s.spmv_format(False)}: iterations "
          f"{info_na['iterations']} / {info['iterations']}", flush=True)
    for d in dev:
        d.free()
    s.close()
    return x_na, info_na, bn, x, info


@pytest.mark.parametrize("first,second", [
    ((60, 50, ox.CG), (20, 16, ox.BICGSTAB)),
    ((40, 40, ox.BICGSTAB), (9, 7, ox.CG)),
])
def test_noalloc_equals_solve_twice_then_smaller(first, second):
    for nx, ny, method in (first, second):
        sc = W.quad_flow(nx, ny, seed=nx + ny, n_steps=1)
        sc.u0 = np.random.default_rng(nx).standard_normal(sc.n)
        x_na, info_na, bn, x, info = _both(sc, method)
        assert np.array_equal(x_na, x)
        for key in ("iterations", "converged", "ret", "error", "bnorm"):
            assert info_na[key] == info[key], (key, info_na[key], info[key])
        assert bn == info["bnorm"] and info["converged"] == 1


def test_noalloc_check_norm_rhs_and_limits():
    """bNorm carried by the caller: a right-hand side 1e-12 of the last one
    is skipped (CheckNormRHS) and leaves x as it was; sizes beyond the fixed
    array and bad arguments fail before any device work."""
    sc = W.quad_flow(12, 10, seed=3, n_steps=1)
    pat, s, ptr, col, a, b, x0 = _assembled(sc, ox.CG)
    dev = [DevArray(v) for v in (ptr, col, a, b * 1e-12, x0)]
    info, bn = ox.solve_crs_noalloc(sc.n, *[d.ptr for d in dev], bnorm=float(np.linalg.norm(b)),
                                    method=ox.CG, tol=1e-10)
    assert info["converged"] == 2 and info["ret"] == 0
    assert np.array_equal(dev[4].get(), x0)
    with pytest.raises(ox.OgsxError, match="maximum of 524288"):
        ox.solve_crs_noalloc(ox.NOALLOC_MAX_ROWS + 1, *[d.ptr for d in dev])
    with pytest.raises(ox.OgsxError, match="null argument"):
        ox.solve_crs_noalloc(sc.n, dev[0].ptr, 0, dev[2].ptr, dev[3].ptr, dev[4].ptr)
    for d in dev:
        d.free()
    s.close()
