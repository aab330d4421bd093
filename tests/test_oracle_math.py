"""The oracle pinned by mathematics and by the reference's own golden values.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a transposed Jacobian, a wrong Gauss weight, a sign or index error)
fails it: exact element integrals from 1-D tensor products or P1 formulas,
invariants (K*1 = 0, Sum dN = 0), brute-force CRS adjacency, dense solves,
and the values held by tests/Matrix/testMatrix.cpp and tests/data/bmskel.
"""
import ctypes as C
import itertools

import numpy as np
import pytest

from ogs5_amd import mesh as M
from ogs5_amd import workloads as W
from oracle import driver

from closed_forms import bmskel_node1

LIQ, HEAT = 1, 3


def _prob(ffi, mesh, **kw):
    kw.setdefault("process", LIQ)
    kw.setdefault("mass_coef", np.ones(mesh.n_elem))
    kw.setdefault("cond", np.ones(mesh.n_elem))
    return ffi.OracleProblem(mesh, **kw)


def _box_exact(coords, h, k, s):
    """Exact bilinear/trilinear Laplace and mass matrices of an axis-aligned
    box element from 1-D linear-element matrices (tensor products)."""
    dim = len(h)
    lo = coords.min(0)
    K1 = [np.array([[1, -1], [-1, 1]]) / hh for hh in h]
    M1 = [np.array([[2, 1], [1, 2]]) * hh / 6 for hh in h]
    idx = np.rint((coords[:, :dim] - lo[:dim]) / np.array(h)).astype(int)  # 0/1 per axis
    n = coords.shape[0]
    Kx = np.zeros((n, n))
    Mx = np.ones((n, n))
    for d in range(dim):
        Md = M1[d][np.ix_(idx[:, d], idx[:, d])]
        Mx = Mx * Md
    for d in range(dim):
        term = np.ones((n, n))
        for e in range(dim):
            mat = K1[e] if e == d else M1[e]
            term = term * mat[np.ix_(idx[:, e], idx[:, e])]
        Kx += term
    return k * Kx, s * Mx


@pytest.mark.parametrize("hx,hy", [(1.0, 1.0), (5.0, 5.0), (2.0, 0.5)])
def test_quad_matrices_closed_form(ffi, hx, hy):
    m = M.quad_mesh(1, 1, hx, hy)
    k, s, area = 3.0e-4, 2.0e-4, 1.7
    p = _prob(ffi, m, mass_coef=[s], cond=[k], geo_area=area)
    Mo, Lo, _, _ = p.element_matrices(0)
    coords = np.stack([m.x, m.y, m.z], 1)[m.conn[0]]
    Ke, Me = _box_exact(coords, (hx, hy), k, s)
    np.testing.assert_allclose(Lo, area * Ke, rtol=1e-13, atol=1e-13 * abs(Ke).max())
    np.testing.assert_allclose(Mo, area * Me, rtol=1e-13, atol=0)
    if hx == hy:  # the classic square stencil K[2/3,-1/6,-1/3,-1/6], M A/36 [4,2,1,2]
        np.testing.assert_allclose(Lo[0] / (area * k), [2 / 3, -1 / 6, -1 / 3, -1 / 6], rtol=1e-13)
        np.testing.assert_allclose(Mo[0] / (area * s * hx * hy), np.array([4, 2, 1, 2]) / 36,
                                   rtol=1e-13)


def test_hex_matrices_closed_form(ffi):
    m = M.hex_mesh(1, 1, 1, 2.0, 1.5, 0.5)
    k, s = 2.5e-5, 1e-4
    p = _prob(ffi, m, mass_coef=[s], cond=[k])
    Mo, Lo, _, _ = p.element_matrices(0)
    coords = np.stack([m.x, m.y, m.z], 1)[m.conn[0]]
    Ke, Me = _box_exact(coords, (2.0, 1.5, 0.5), k, s)
    np.testing.assert_allclose(Lo, Ke, rtol=1e-13, atol=1e-13 * abs(Ke).max())
    np.testing.assert_allclose(Mo, Me, rtol=1e-13)
    # unit cube: diag h/3, edge 0, face-diagonal -h/12, body-diagonal -h/12
    u = M.hex_mesh(1, 1, 1)
    Lu = _prob(ffi, u, cond=[1.0]).element_matrices(0)[1]
    assert abs(Lu[0, 0] - 1 / 3) < 1e-14 and abs(Lu[0, 1]) < 1e-15
    assert abs(Lu[0, 2] + 1 / 12) < 1e-14 and abs(Lu[0, 6] + 1 / 12) < 1e-14


def test_hex_full_tensor_and_distortion(ffi):
    """Anisotropic tensor on a distorted hex: symmetry, K*1 = 0, and the
    energy u^T K u of a linear field u = g.x equals V g^T T g."""
    rng = np.random.default_rng(5)
    m = M.hex_mesh(1, 1, 1)
    m.x = m.x + 0.1 * rng.standard_normal(8) * 0  # keep faces planar for exactness
    A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
    X = np.stack([m.x, m.y, m.z], 1) @ A.T  # affine map keeps trilinear exact
    m.x, m.y, m.z = X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy()
    T = np.array([[2.0, 0.3, 0.1], [0.3, 1.0, -0.2], [0.1, -0.2, 0.5]])
    p = _prob(ffi, m, cond=T.ravel(), cond_stride=9)
    L = p.element_matrices(0)[1]
    np.testing.assert_allclose(L, L.T, atol=1e-14)
    np.testing.assert_allclose(L.sum(1), 0, atol=1e-13)
    g = np.array([0.3, -1.1, 0.7])
    u = X[m.conn[0]] @ g
    vol = abs(np.linalg.det(A))
    assert abs(u @ L @ u - vol * g @ T @ g) < 1e-12 * abs(vol * g @ T @ g)


def _tet_exact(X, k, s, v):
    B = np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]])
    V = abs(np.linalg.det(B)) / 6
    G = np.linalg.inv(B)  # columns: gradients of phi_1..3
    grads = np.vstack([-G.sum(1), G.T])  # [4,3]
    K = k * V * grads @ grads.T
    Mm = s * V / 20 * (np.ones((4, 4)) + np.eye(4))
    A = V / 4 * np.outer(np.ones(4), grads @ v)
    return K, Mm, A


def test_tet_matrices_exact_p1(ffi):
    rng = np.random.default_rng(3)
    X = np.array([[0, 0, 0], [1.2, 0.1, 0], [0.2, 0.9, 0.1], [0.1, 0.3, 1.1]]) + 0.05 * rng.random((4, 3))
    m = M.Mesh(M.TET4, 3, X[:, 0].copy(), X[:, 1].copy(), X[:, 2].copy(),
               np.array([[0, 1, 2, 3]], np.int64), (1, 1, 1), (1, 1, 1), np.zeros(1, np.int64))
    v = np.array([0.4, -0.2, 0.9])
    vel = np.repeat(v[:, None], 5, axis=1)[None]  # [E][3][5]
    p = _prob(ffi, m, process=HEAT, mass_coef=[2.0], cond=[0.7], adv_coef=1.3, vel=vel)
    Mo, Lo, Ao, _ = p.element_matrices(0)
    Ke, Me, Ae = _tet_exact(X, 0.7, 2.0, 1.3 * v)
    np.testing.assert_allclose(Lo, Ke, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(Mo, Me, rtol=1e-13)
    np.testing.assert_allclose(Ao, Ae, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(Ao.sum(1), 0, atol=1e-14)


def test_line_matrices(ffi):
    m = M.line_mesh(1, 2.0)
    p = _prob(ffi, m, mass_coef=[3.0], cond=[0.5], geo_area=4.0)
    Mo, Lo, _, _ = p.element_matrices(0)
    np.testing.assert_allclose(Lo, 4.0 * 0.5 / 2.0 * np.array([[1, -1], [-1, 1]]), rtol=1e-14)
    np.testing.assert_allclose(Mo, 4.0 * 3.0 * 2.0 / 6 * np.array([[2, 1], [1, 2]]), rtol=1e-13)


def test_gravity_vector_unit_cube(ffi):
    m = M.hex_mesh(1, 1, 1)
    p = _prob(ffi, m, cond=[2.0], rho_g=9.81e3)
    G = p.element_matrices(0)[3]
    top = m.z[m.conn[0]] > 0.5
    # -rho_g k * int dN_i/dz dV = -rho_g k (+-1/4)
    np.testing.assert_allclose(G, np.where(top, -1, 1) * 9.81e3 * 2.0 / 4, rtol=1e-13)


def test_time_discretisation_identities(ffi):
    m = M.quad_mesh(1, 1, 3.0, 2.0)
    u = np.array([1.0, 2.0, -0.5, 0.25])
    dt = 7.0
    for proc, theta in [(LIQ, 1.0), (LIQ, 0.5), (HEAT, 0.5), (HEAT, 1.0)]:
        p = _prob(ffi, m, process=proc, mass_coef=[3.0], cond=[0.4], dt=dt, theta=theta, u_old=u)
        Mo, Lo, _, _ = p.element_matrices(0)
        lhs, rhs = p.element_system(0)
        ul = u[m.conn[0]]  # local node order of the element
        # LHS*u - RHS = (M/dt + theta K) u - (M/dt - (1-theta) K) u = K u
        np.testing.assert_allclose(lhs @ ul - rhs, Lo @ ul, rtol=1e-12, atol=1e-12)
        # constant field: K*1 = 0, so RHS sums to the total storage / dt
        p.set(u_old=np.full(4, 2.0))
        np.testing.assert_allclose(p.element_system(0)[1].sum(), 2.0 * Mo.sum() / dt, rtol=1e-13)


@pytest.mark.parametrize("theta", [1.0, 0.5, 0.25, 0.0])
@pytest.mark.parametrize("proc", [LIQ, HEAT])
def test_theta_split_pinned(ffi, proc, theta):
    """The theta split itself, against closed-form element matrices:
    LHS = M/dt + th*(K+A) and RHS = (M/dt - (1-th)*(K+A)) u_n, with
    th = nls_relaxation for LIQUID_FLOW (0 taken as 1,
    fem_ele_std.cpp:7715-7720, :7841-7880) and th = ls_theta for
    HEAT_TRANSPORT (:8509, :8630-8694).  theta = 0.25 fails a swap of theta
    and 1 - theta; theta = 1 fails explicit in place of implicit Euler."""
    hx, hy, k, s, dt, area = 3.0, 2.0, 0.4, 3.0, 7.0, 1.3
    m = M.quad_mesh(1, 1, hx, hy)
    u = np.array([1.0, 2.0, -0.5, 0.25])
    coords = np.stack([m.x, m.y, m.z], 1)[m.conn[0]]
    Ke, Me = _box_exact(coords, (hx, hy), k, s)
    Ke, Me = area * Ke, area * Me
    kw = {}
    Ae = np.zeros_like(Ke)
    if proc == HEAT:
        # constant velocity: A_ij = int N_i (c v . grad N_j), exact on the
        # bilinear element as the 1-D tensor products of N and dN/dx
        v, c = np.array([0.7, -0.3, 0.0]), 2.0
        idx = np.rint((coords[:, :2] - coords[:, :2].min(0)) / (hx, hy)).astype(int)
        N1 = [np.array([[2, 1], [1, 2]]) * h / 6 for h in (hx, hy)]  # int N_a N_b
        D1 = [np.array([[-1, 1], [-1, 1]]) / 2.0 for _ in (hx, hy)]  # int N_a dN_b
        for d in range(2):
            term = np.ones((4, 4))
            for e in range(2):
                mat = D1[e] if e == d else N1[e]
                term = term * mat[np.ix_(idx[:, e], idx[:, e])]
            Ae += c * v[d] * term
        Ae *= area
        kw = dict(adv_coef=c, vel=np.repeat(v[:, None], 9, axis=1)[None])
    p = _prob(ffi, m, process=proc, mass_coef=[s], cond=[k], dt=dt, theta=theta, u_old=u,
              geo_area=area, **kw)
    Mo, Lo, Ao, _ = p.element_matrices(0)
    np.testing.assert_allclose(Ao, Ae, rtol=1e-12, atol=1e-13 * abs(Ke).max())
    lhs, rhs = p.element_system(0)
    ul = u[m.conn[0]]
    th = 1.0 if (proc == LIQ and theta == 0.0) else theta
    KA = Ke + Ae
    scale = abs(Me / dt).max() + abs(KA).max()
    np.testing.assert_allclose(lhs, Me / dt + th * KA, rtol=0, atol=1e-13 * scale)
    np.testing.assert_allclose(rhs, (Me / dt - (1.0 - th) * KA) @ ul, rtol=0,
                               atol=1e-13 * scale * abs(ul).max())


def test_implicit_euler_huge_step_is_the_steady_solve(ffi):
    """theta = 1 with dt -> infinity: one time step from any u_n lands on the
    steady solution of K u = f (dense solve of the closed-form system).  The
    explicit step (theta and 1 - theta swapped) would return M/dt-weighted
    u_n - K u_n instead."""
    sc = W.quad_flow(6, 5, seed=3, n_steps=1)
    sc.steps = [W.Step(1e14, 1.0)]
    sc.u0 = np.random.default_rng(1).standard_normal(sc.n) * 5.0
    r = driver.run_oracle(sc, ffi, method=2, symmetric=False, tol=1e-14)
    steady = W.quad_flow(6, 5, seed=3, n_steps=1)
    steady.steps = [W.Step(1.0, 1.0, True)]
    rs = driver.run_oracle(steady, ffi, method=2, symmetric=False, tol=1e-14, keep_steps=True)
    eq = rs["eqs"]
    a, b = rs["systems"][0]
    eq.a[:] = a
    x_dense = np.linalg.solve(eq.dense(), b)
    assert np.linalg.norm(r["x"] - x_dense) / np.linalg.norm(x_dense) < 1e-9


def test_reference_dense_multi_golden(ffi):
    """tests/Matrix/testMatrix.cpp: [[1,2,3],[4,5,6],[7,8,9]]*[1,2,3] = [14,32,50];
    after m = 1; m*=10; m/=2; m+=5 every row of m*1 is 30."""
    L = ffi.lib()
    m = np.arange(1.0, 10.0)
    x = np.array([1.0, 2.0, 3.0])
    b = np.zeros(3)
    L.orc_dense_multi(m.ctypes.data_as(C.c_void_p), 3, 3, x.ctypes.data_as(C.c_void_p),
                      b.ctypes.data_as(C.c_void_p), 1.0)
    assert list(b) == [14.0, 32.0, 50.0]
    m = np.full(9, (1.0 * 10.0) / 2.0 + 5.0)
    b[:] = 0
    ones = np.ones(3)
    L.orc_dense_multi(m.ctypes.data_as(C.c_void_p), 3, 3, ones.ctypes.data_as(C.c_void_p),
                      b.ctypes.data_as(C.c_void_p), 1.0)
    assert list(b) == [30.0, 30.0, 30.0]


def _brute_pattern(n, conn):
    rows = [set() for _ in range(n)]
    for e in conn:
        for a in e:
            rows[a].update(int(c) for c in e)
    ptr = np.zeros(n + 1, np.int64)
    col = []
    diag = np.full(n, -1, np.int64)
    for i in range(n):
        cs = sorted(rows[i])
        for c in cs:
            if c == i:
                diag[i] = len(col)
            col.append(c)
        ptr[i + 1] = len(col)
    return ptr, np.array(col, np.int64), diag


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pattern_brute_force_random_meshes(ffi, seed):
    rng = np.random.default_rng(seed)
    n = 40
    for npe in (2, 4, 8):
        conn = np.array([rng.choice(n - 3, npe, replace=False) for _ in range(25)], np.int64)
        got = ffi.pattern(n, conn)
        exp = _brute_pattern(n, conn)
        for g, e in zip(got, exp):
            np.testing.assert_array_equal(g, e)


@pytest.mark.parametrize("mk,nnz", [
    (lambda: M.quad_mesh(200, 200), (3 * 201 - 2) ** 2),
    (lambda: M.hex_mesh(9, 9, 9), (3 * 10 - 2) ** 3),
    (lambda: M.tet_mesh(7, 7, 7), 8 ** 3 + 2 * (3 * 64 * 7 + 3 * 8 * 49 + 343)),
])
def test_pattern_closed_form_nnz(ffi, mk, nnz):
    m = mk()
    ptr, col, diag = ffi.pattern(m.n_nodes, m.conn)
    assert ptr[-1] == nnz
    assert np.all(np.diff(ptr) > 0)
    for i in range(0, m.n_nodes, max(1, m.n_nodes // 50)):
        seg = col[ptr[i]:ptr[i + 1]]
        assert np.all(np.diff(seg) > 0) and col[diag[i]] == i


def _small_system(ffi, symmetric, method, tol=1e-13):
    sc = W.quad_flow(10, 8, seed=11, n_steps=1)
    r = driver.run_oracle(sc, ffi, method=method, symmetric=symmetric, keep_steps=True, tol=tol)
    eq = r["eqs"]
    a, b = r["systems"][0]
    eq.a[:] = a
    return sc, r, eq.dense(), b


def test_dirichlet_row_only_structure(ffi):
    sc, r, A, b = _small_system(ffi, False, 2)
    for i, g in zip(sc.bc_idx, sc.bc_val):
        row = A[i].copy()
        assert row[i] != 0 and np.count_nonzero(np.delete(row, i)) == 0
        assert b[i] == A[i, i] * g
    # columns of Dirichlet nodes survive in other rows (nonsymmetric)
    i = sc.bc_idx[0]
    assert np.count_nonzero(A[:, i]) > 1 and not np.allclose(A, A.T)


def test_dirichlet_symmetric_same_solution(ffi):
    sc, r1, A1, b1 = _small_system(ffi, False, 2)
    sc, r2, A2, b2 = _small_system(ffi, True, 5)
    np.testing.assert_allclose(A2, A2.T, atol=1e-18)
    x1 = np.linalg.solve(A1, b1)
    x2 = np.linalg.solve(A2, b2)
    np.testing.assert_allclose(x1, x2, rtol=1e-11, atol=1e-11 * abs(x1).max())


@pytest.mark.parametrize("method,symmetric", [(2, False), (5, True), (2, True)])
def test_krylov_matches_dense_solve(ffi, method, symmetric):
    sc, r, A, b = _small_system(ffi, symmetric, method, tol=1e-13)
    x = np.linalg.solve(A, b)
    got = r["steps"][0]
    assert np.linalg.norm(got - x) / np.linalg.norm(x) < 1e-10
    assert r["info"][0]["error"] < 1e-13


def test_check_norm_rhs_and_max_iter(ffi):
    sc = W.quad_flow(10, 8, seed=11, n_steps=1)
    ptr, col, diag = ffi.pattern(sc.n, sc.mesh.conn)
    r = driver.run_oracle(sc, ffi, method=5, symmetric=True, keep_steps=True, tol=1e-10)
    a, b = r["systems"][0]
    eq = ffi.OracleEqs(ptr, col, diag)
    eq.config(5, 1, 1e-10, 3)
    eq.a[:] = a
    eq.b[:] = b
    out = eq.solve()  # CG: returns iter <= max_iter, here 0 (not converged)
    assert out["ret"] == 0 and out["iterations"] == 4
    # a right-hand side 1e-12 of the previous one is skipped by CheckNormRHS
    eq.b[:] = b * 1e-12
    x_before = eq.x.copy()
    out = eq.solve()
    assert out["ret"] == 0 and np.array_equal(eq.x, x_before)
    eq2 = ffi.OracleEqs(ptr, col, diag)
    eq2.config(2, 1, 1e-10, 3)
    eq2.a[:] = a
    eq2.b[:] = b
    out = eq2.solve()  # BiCGStab returns the loop counter: max_iter + 1
    assert out["ret"] == 4


def test_bmskel_golden(ffi):
    """tests/data/bmskel (testBmSkel.cxx:257-291): T(POINT0) = 274 after the
    one time step, solved with BiCGStab + Jacobi on row-only Dirichlet."""
    sc = W.bmskel()
    r = driver.run_oracle(sc, ffi, method=2, symmetric=False)
    assert r["x"][0] == 274.0
    assert abs(r["x"][1] - bmskel_node1()) <= 1e-12 * abs(bmskel_node1())



def test_write_bin_layout(ffi, tmp_path):
    sc, r, A, b = _small_system(ffi, False, 2)
    eq = r["eqs"]
    p = tmp_path / "eq.bin"
    assert eq.write_bin(p) == 0
    n, ptr, col, a, bb = ffi.read_bin(p)
    assert n == sc.n and np.array_equal(ptr, r["ptr"]) and np.array_equal(col, r["col"])
    np.testing.assert_array_equal(bb, eq.b)


def test_reference_bicgstab_breakdown_on_dirichlet_only_rhs(ffi):
    """A property of the reference algorithm that the tests rely on: with
    row-only Dirichlet elimination, x0 = 0 and b nonzero only on Dirichlet
    rows (a steady solve without sources), one Jacobi-preconditioned step
    makes s and t vanish on those rows, so (r0, r) = 0 exactly at iteration 2
    and Linear_EQS::BiCGStab returns 0 with an unconverged x.  CG on the
    symmetric elimination solves the same system."""
    sc = W.hex_flow(6, seed=6, n_transient=0)
    r = driver.run_oracle(sc, ffi, method=2, symmetric=False, tol=1e-13, keep_steps=True)
    assert r["info"][0]["ret"] == 0 and r["info"][0]["error"] > 0.1
    c = driver.run_oracle(sc, ffi, method=5, symmetric=True, tol=1e-13)
    assert c["info"][0]["ret"] == 1 and c["info"][0]["error"] < 1e-13


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_jds_storage_sums_like_crs(ffi, seed):
    """The reference's JDS storage (rows sorted by decreasing length, jagged
    columns) sums every row in the same order as CRS, so its SpMV equals the
    CRS SpMV bit for bit: the device path, which keeps CRS order, serves
    decks with storage 4 unchanged."""
    rng = np.random.default_rng(seed)
    m = [M.tet_mesh(3, 2, 2), M.quad_mesh(5, 4), M.hex_mesh(2, 3, 2)][seed]
    ptr, col, diag = ffi.pattern(m.n_nodes, m.conn)
    a = rng.standard_normal(int(ptr[-1]))
    x = rng.standard_normal(m.n_nodes)
    eq = ffi.OracleEqs(ptr, col, diag)
    eq.a[:] = a
    y_crs = eq.spmv(x)
    y_jds, lay = ffi.jds_spmv(ptr, col, a, x)
    assert np.array_equal(y_jds, y_crs)
    lens = np.diff(ptr)[lay["n2o"]]
    assert np.all(np.diff(lens) <= 0)  # decreasing row length
    for L_ in np.unique(lens):  # equal lengths keep their order (stable)
        rows = lay["n2o"][lens == L_]
        assert np.all(np.diff(rows) > 0)
    assert np.array_equal(lay["ncol"], [(np.diff(ptr) > k).sum() for k in range(lens[0])])


def test_groundwater_flow_is_the_head_form(ffi):
    """GROUNDWATER_FLOW: Laplace K*tuf (no viscosity), mass = storage, no
    gravity term even with a density set, the same parabolic step; its
    Gauss-point velocity is -K grad h."""
    rng = np.random.default_rng(4)
    m = M.hex_mesh(2, 2, 2)
    T = np.array([[2.0, 0.3, 0.1], [0.2, 1.0, -0.2], [0.1, -0.1, 0.5]]) * 1e-5
    kw = dict(mass_coef=np.full(m.n_elem, 3e-4), cond=np.tile(T.ravel(), m.n_elem),
              cond_stride=9, tuf=86400.0, dt=2.0, theta=1.0, u_old=rng.standard_normal(m.n_nodes))
    gw = ffi.OracleProblem(m, process=2, rho_g=9.81e3, mu=7.0, **kw)
    liq = ffi.OracleProblem(m, process=LIQ, rho_g=0.0, mu=1.0, **kw)
    for e in range(m.n_elem):
        Mg, Lg, _, Gg = gw.element_matrices(e)
        Ml, Ll, _, _ = liq.element_matrices(e)
        assert np.array_equal(Mg, Ml) and np.array_equal(Lg, Ll)  # tuf*K/1 == K*tuf*1
        assert not Gg.any()
        assert all(np.array_equal(a, b) for a, b in zip(gw.element_system(e),
                                                        liq.element_system(e)))
    g = np.array([0.4, -1.3, 0.8])
    h = np.stack([m.x, m.y, m.z], 1) @ g
    v = gw.gauss_velocity(h)
    assert np.abs(v - (-(T @ g))[None, :, None]).max() < 1e-12 * np.abs(T @ g).max()
