"""HIP path vs the oracle, through the C ABI, on the same seeded inputs.

Bars: CRS index arrays bit-exact; assembled FP64 entries within 1e-12 of
the oracle's, relative to the largest entry of their row (entries that cancel
to ~0 in the global sum are compared on the row's scale); right-hand sides
within 1e-12 relative to max|b|; solutions within 1e-10 relative L2 of a
tight-tolerance oracle solve.
"""
import numpy as np
import pytest

import ogs5_amd as ox
from ogs5_amd import workloads as W
from ogs5_amd.fields import lognormal_conductivity
from ogs5_amd.mesh import hex_mesh, line_mesh, quad_mesh
from oracle import driver

pytestmark = pytest.mark.gpu

ENTRY_TOL = 1e-12
SOL_TOL = 1e-10
# Relative L2 error bounds of the full-size solutions (FP64 and FP32-inside-
# FP64 refinement, both stopped at a 1e-10 relative residual) against a
# 1e-13 FP64 solve of the same system.  Measured: 128^3 hexes, sigma_lnK = 2:
# FP64 6.7e-9, IR 9.4e-9 (the condition number turns a 1e-10 residual into an
# error near 1e-8); 96^3 x 6 tets: FP64 1.3e-11 and 3.2e-11 (two builds whose
# element kernels contract different products to FMA: the stopping residual,
# 2.3e-11 against 5.7e-11, moves with the last bits of A), IR 5.4e-14.  The
# tet bound is the stopping tolerance, the hex bound twice the measured error.
HEX_ERR_BOUND = 2e-8
TET_ERR_BOUND = 1e-10


def _scenarios():
    out = {
        "quad": W.quad_flow(20, 16, seed=5, n_steps=3),
        "hex": W.hex_flow(6, seed=6, n_transient=2),
        "tet": W.tet_heat(5, seed=7, n_steps=3),
        "bmskel": W.bmskel(),
    }
    # full anisotropic tensor, distorted hexes, gravity on
    sc = W.hex_flow(5, seed=8, n_transient=2, steady=False)
    rng = np.random.default_rng(8)
    m = sc.mesh
    m.x = m.x + 0.15 * rng.uniform(-1, 1, m.n_nodes)
    m.y = m.y + 0.15 * rng.uniform(-1, 1, m.n_nodes)
    T = np.array([[2.0, 0.3, 0.1], [0.3, 1.0, -0.2], [0.1, -0.2, 0.5]]) * 1e-4
    sc.cond = np.tile(T.ravel(), (m.n_elem, 1)) * sc.cond[:, None] / 1e-4
    sc.cond_stride = 9
    sc.rho_g = 1000.0 * 9.81
    out["hex_aniso_grav"] = sc
    # GROUNDWATER_FLOW (HEAD form): viscosity and density set but not used
    gw = W.quad_flow(20, 16, seed=9, n_steps=3)
    gw.process, gw.tuf, gw.mu, gw.rho_g = ox.GROUNDWATER_FLOW, 86400.0, 5.0, 9.81e3
    gw.steps = [W.Step(st.dt / 86400.0, st.theta) for st in gw.steps]
    out["quad_gw"] = gw
    # other $ELE_GAUSS_POINTS: 1, 2 and 4 points per direction (4^3 = 64 on
    # hexes, the largest Gauss table)
    for name, sc2, ng in (("quad_ng1", W.quad_flow(10, 8, seed=11, n_steps=2), 1),
                          ("hex_ng2", W.hex_flow(5, seed=12, n_transient=1), 2),
                          ("hex_ng4", W.hex_flow(4, seed=13, n_transient=1, steady=False), 4),
                          ("quad_ng4", W.quad_flow(12, 9, seed=14, n_steps=2), 4)):
        sc2.n_gauss = ng
        out[name] = sc2
    out["hex_ng4"].cond = np.tile((T / 1e-4).ravel(), (out["hex_ng4"].mesh.n_elem, 1)) * \
        out["hex_ng4"].cond[:, None]
    out["hex_ng4"].cond_stride = 9
    # 3-D flow with a source and a nonzero start: the right-hand side does
    # not live on the Dirichlet rows alone, so the reference's BiCGStab does
    # not break down and hexes get BiCGStab parity too
    hs = W.hex_flow(7, seed=19, n_transient=2, steady=False)
    hs.st_idx = np.array([hs.mesh.node_index(3, 4, 3)], np.int64)
    hs.st_val = np.array([-2e-4])
    hs.u0 = np.random.default_rng(19).uniform(0.0, 1.0, hs.n)
    out["hex_sink"] = hs
    # geo_area != 1 multiplies |J| of elements of dimension < 3
    # (SetMaterial, fem_ele_std.cpp:948-956): a 2-D flow with a full tensor
    # and gravity on the second axis, a 2-D heat transport with advection,
    # and 1-D flow with gravity
    qg = W.quad_flow(14, 11, seed=15, n_steps=2)
    T2 = np.array([[1.5, 0.4], [0.3, 0.8]])
    qg.cond = np.tile(T2.ravel(), (qg.mesh.n_elem, 1)) * qg.cond[:, None]
    qg.cond_stride = 4
    qg.geo_area, qg.rho_g, qg.mu = 1.7, 9.81e3, 1.3e-3
    out["quad_geo"] = qg
    rng = np.random.default_rng(16)
    mq = quad_mesh(12, 10, 30.0, 25.0)
    Eq = mq.n_elem
    hq = W.Scenario("quad_heat_geo", mq, ox.HEAT_TRANSPORT, np.full(Eq, 2.5e6),
                    np.full(Eq, 2.0), 1, mq.nodes_where(0, 0.0),
                    np.full(mq.nodes_where(0, 0.0).size, 60.0), np.zeros(0, np.int64), np.zeros(0),
                    np.full(mq.n_nodes, 10.0), [W.Step(3.0e4, 0.5)] * 3, geo_area=1.7,
                    adv_coef=4.2e6, vel=np.tile(np.array([2e-6, -5e-7, 0.0])[:, None], (Eq, 1, 9))
                    * rng.uniform(0.5, 1.5, (Eq, 1, 1)))
    out["quad_heat_geo"] = hq
    ml = line_mesh(30, 45.0)
    El = ml.n_elem
    out["line_geo"] = W.Scenario("line_geo", ml, ox.LIQUID_FLOW, np.full(El, 1e-4),
                                 1e-4 * np.exp(rng.normal(0.0, 1.0, El)), 1,
                                 np.array([0, ml.n_nodes - 1], np.int64), np.array([5.0, 1.0]),
                                 np.array([11], np.int64), np.array([-1e-6]),
                                 np.zeros(ml.n_nodes), [W.Step(600.0, 1.0)] * 3,
                                 geo_area=1.7, mu=1e-3, rho_g=9.81e3)
    return out


SCEN = _scenarios()


def _row_scaled_close(a, a_ref, ptr, tol):
    n = ptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(ptr))
    rowmax = np.zeros(n)
    np.maximum.at(rowmax, rows, np.abs(a_ref))
    err = np.abs(a - a_ref) / np.maximum(rowmax[rows], 1e-300)
    return err.max()


def _assemble_both(sc, bc_mode, step=0):
    """Assembles one step (with ST and, unless bc_mode is None, Dirichlet) on
    the GPU and in the oracle; returns (A_gpu, b_gpu, A_ref, b_ref, ptr)."""
    from oracle import ffi
    pat, s = W.make_system(sc)
    st = sc.steps[step]
    s.set_vector(ox.VEC_UOLD, sc.u0)
    s.initialize()
    s.assemble(st.dt, st.theta, st.steady)
    if sc.st_idx.size:
        s.add_rhs(sc.st_idx, sc.st_val)
    if bc_mode is not None:
        s.set_dirichlet(sc.bc_idx, sc.bc_val, bc_mode)
    A, b = s.get_matrix(), s.get_vector(ox.VEC_B)
    ptr, col, diag = ffi.pattern(sc.n, sc.mesh.conn)
    eq = ffi.OracleEqs(ptr, col, diag)
    prob = ffi.OracleProblem(sc.mesh, process=sc.process, mass_coef=sc.mass_coef, cond=sc.cond,
                             cond_stride=sc.cond_stride, n_gauss=sc.n_gauss, tuf=sc.tuf,
                             mu=sc.mu, geo_area=sc.geo_area, adv_coef=sc.adv_coef, vel=sc.vel,
                             rho_g=sc.rho_g, dt=st.dt, theta=st.theta, steady=st.steady,
                             u_old=sc.u0)
    eq.initialize()
    eq.assemble(prob)
    if sc.st_idx.size:
        eq.add_rhs(sc.st_idx, sc.st_val)
    if bc_mode is not None:
        eq.set_known_x(sc.bc_idx, sc.bc_val, symmetric=(bc_mode == ox.DIRICHLET_SYMMETRIC))
    p_exp = pat.export()
    assert np.array_equal(p_exp[0], ptr) and np.array_equal(p_exp[1], col)
    return A, b, eq.a.copy(), eq.b.copy(), ptr, s, eq


@pytest.mark.parametrize("name", list(SCEN))
@pytest.mark.parametrize("bc", [None, ox.DIRICHLET_ROW, ox.DIRICHLET_SYMMETRIC])
def test_assembly_parity(name, bc):
    sc = SCEN[name]
    A, b, Ar, br, ptr, _, _ = _assemble_both(sc, bc, step=min(1, len(sc.steps) - 1))
    print(f"[assembly] {name} bc={bc} n={sc.n} nnz={ptr[-1]}", flush=True)
    assert _row_scaled_close(A, Ar, ptr, ENTRY_TOL) < ENTRY_TOL
    assert np.abs(b - br).max() <= ENTRY_TOL * max(np.abs(br).max(), 1e-300)


def _star_mesh(m):
    """Line elements from a hub node 0 to each of nodes 1..m, plus the chain
    1-2-...-m: the hub's CRS row holds m + 1 entries."""
    from ogs5_amd import mesh as M
    k = np.arange(1, m + 1)
    x = np.r_[0.0, 0.1 * k + 0.013 * np.sin(k)]
    conn = np.array([(0, i) for i in k] + [(i, i + 1) for i in range(1, m)], dtype=np.int64)
    z = np.zeros_like(x)
    return M.Mesh(M.LINE2, 1, x, z.copy(), z.copy(), conn, (len(conn),), (0.1,),
                  np.arange(len(conn), dtype=np.int64))


@pytest.mark.parametrize("bc", [None, ox.DIRICHLET_ROW, ox.DIRICHLET_SYMMETRIC])
@pytest.mark.parametrize("process", [ox.LIQUID_FLOW, ox.HEAT_TRANSPORT])
def test_assembly_parity_long_rows(process, bc):
    """Rows longer than the two-pass assembly's 32-entry row accumulator
    (a hub node in 40 line elements) take the coloured scatter: A and b
    against the oracle, with gravity (flow) and advection (heat).  The
    Dirichlet node (the chain's last node) sits at column 40 of the hub's
    row, so the symmetric elimination lifts it from the row's second chunk
    of 32 entries."""
    m = _star_mesh(40)
    rng = np.random.default_rng(9)
    E = m.n_elem
    flow = process == ox.LIQUID_FLOW
    vel = None if flow else rng.uniform(-1e-6, 1e-6, (E, 3, 3))
    sc = W.Scenario("star", m, process, np.full(E, 1e-4), 1e-5 * np.exp(rng.normal(0.0, 1.0, E)), 1,
                    np.array([m.n_nodes - 1], dtype=np.int64), np.array([1.0]),
                    np.zeros(0, np.int64), np.zeros(0), rng.standard_normal(m.n_nodes),
                    [W.Step(10.0, 1.0 if flow else 0.5)], mu=1e-3 if flow else 1.0,
                    rho_g=9810.0 if flow else 0.0, adv_coef=0.0 if flow else 4.2e6, vel=vel)
    A, b, Ar, br, ptr, s, eq = _assemble_both(sc, bc)
    assert np.diff(ptr).max() > 32
    assert _row_scaled_close(A, Ar, ptr, ENTRY_TOL) < ENTRY_TOL
    assert np.abs(b - br).max() <= ENTRY_TOL * max(np.abs(br).max(), 1e-300)
    s.close()


@pytest.mark.parametrize("lpr", [1, 2, 3, 5, 6, 7, 9, 10, 11, 12, 13, 14, 17, 18, 21, 22, 19, 20,
                                 4, 8, 16, 32, 64])
def test_spmv_parity_all_variants(lpr):
    sc = SCEN["hex"]
    A, b, Ar, br, ptr, s, eq = _assemble_both(sc, ox.DIRICHLET_ROW)
    s.set_spmv_variant(lpr)
    x = np.random.default_rng(lpr).standard_normal(sc.n)
    y = s.spmv(x)
    assert s.spmv_variant() == lpr
    eq.a[:] = A  # same matrix on both sides
    yr = eq.spmv(x)
    if lpr in (1, 2, 3, 5, 6, 7, 9, 10, 11, 12, 13, 14, 17, 18, 21, 22, 19, 20):
        # CRS-order summation over rounded products: the reference's multiVec bits
        assert np.array_equal(y, yr)
        return
    rows = np.repeat(np.arange(sc.n), np.diff(ptr))
    scale = np.zeros(sc.n)
    np.add.at(scale, rows, np.abs(A * x[eq.col]))
    assert np.all(np.abs(y - yr) <= 1e-14 * scale + 1e-300)


@pytest.mark.parametrize("name", ["quad", "tet", "hex_aniso_grav", "bmskel"])
@pytest.mark.parametrize("fmt", [13, 14, 17, 18, 21, 22])
def test_spmv_stencil_dictionary_bit_exact(name, fmt):
    """The dictionary formats on 2-D, tet and small meshes (padding rows in
    the last slice): the reference's multiVec bits."""
    sc = SCEN[name]
    A, b, Ar, br, ptr, s, eq = _assemble_both(sc, ox.DIRICHLET_SYMMETRIC)
    s.set_spmv_variant(fmt)
    x = np.random.default_rng(fmt).standard_normal(sc.n)
    y = s.spmv(x)
    eq.a[:] = A
    assert np.array_equal(y, eq.spmv(x))


@pytest.mark.parametrize("name", ["quad", "hex", "tet", "hex_aniso_grav", "bmskel", "quad_ng4"])
@pytest.mark.parametrize("fmt", [19, 20])
def test_spmv_slice_triples_bit_exact(name, fmt):
    """Slice triples (one stencil per 64-row slice, lane-shifted gathers,
    zero entries where a row lacks a column of its slice's stencil) on 2-D,
    3-D and 1-D meshes, including a last slice with padding rows and slices
    whose gathers are clamped at the first and last rows: the reference's
    multiVec bits."""
    sc = SCEN[name]
    A, b, Ar, br, ptr, s, eq = _assemble_both(sc, ox.DIRICHLET_SYMMETRIC)
    s.set_spmv_variant(fmt)
    eq.a[:] = A
    for seed in range(3):
        x = np.random.default_rng(100 * fmt + seed).standard_normal(sc.n)
        assert np.array_equal(s.spmv(x), eq.spmv(x))
    # A changed: the value copy follows it in the triple layout
    s.set_spmv_variant(13 if fmt == 19 else 5)
    x = np.random.default_rng(7).standard_normal(sc.n)
    y13 = s.spmv(x)
    s.set_spmv_variant(fmt)
    assert np.array_equal(s.spmv(x), y13)


@pytest.mark.parametrize("prec", [ox.FP64, ox.FP32_IR])
@pytest.mark.parametrize("fmt", [19, 20])
def test_solve_with_slice_triples(prec, fmt):
    """CG solves with the triple formats forced (both precisions of an IR
    system use them) reach the oracle's solution chain."""
    sc = SCEN["hex"]
    ref = _tight_reference(sc)
    pat, s = W.make_system(sc, prec)
    s.set_spmv_variant(fmt)
    res = W.run_gpu(sc, method=ox.CG, dirichlet=ox.DIRICHLET_SYMMETRIC, precision=prec,
                    tol=1e-13, keep_steps=True, system=(pat, s))
    assert s.spmv_variant() == fmt
    for x, xr in zip(res["steps"], ref["steps"]):
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < SOL_TOL
    s.close()


def test_symmetric_diagonal_fp64_copy_identical_to_crs():
    """An FP64 system that is symmetric by construction on a banded pattern
    holds its matrix as the symmetric diagonal copy (format 25), summed in
    CRS order with the rows on the SELL-64 threads: solves and the
    stand-alone SpMV equal the plain CRS format's bit for bit (quads: 4
    offsets, hexes: 13)."""
    for name in ("quad", "hex"):
        sc = SCEN[name]
        v = np.random.default_rng(5).standard_normal(sc.n)
        out = []
        for forced in (None, 9):
            pat, s = W.make_system(sc, ox.FP64)
            if forced is not None:
                s.set_spmv_variant(forced)
            res = W.run_gpu(sc, method=ox.CG, dirichlet=ox.DIRICHLET_SYMMETRIC, keep_steps=True,
                            system=(pat, s))
            assert s.spmv_format(False) == (25 if forced is None else 9), name
            out.append((res, s.spmv(v)))
            s.close()
        (dia, y_dia), (crs, y_crs) = out
        assert [i["iterations"] for i in dia["info"]] == [i["iterations"] for i in crs["info"]]
        for x, xr in zip(dia["steps"], crs["steps"]):
            assert np.array_equal(x, xr), name
        assert np.array_equal(y_dia, y_crs), name


def test_symmetric_diagonal_fp32_copy():
    """An FP32-IR system whose matrix is symmetric by construction, on a
    banded pattern, runs its inner FP32 solve on the symmetric diagonal copy
    (format 25: the diagonal and the upper diagonals, each lower entry read
    from its mirror); the refinement still reaches the oracle's solutions.
    With row-only Dirichlet rows (a nonsymmetric matrix) the FP32 matrix
    keeps a tuned full format."""
    for name in ("quad", "hex"):
        sc = SCEN[name]
        ref = _tight_reference(sc)
        pat, s = W.make_system(sc, ox.FP32_IR)
        res = W.run_gpu(sc, method=ox.CG, dirichlet=ox.DIRICHLET_SYMMETRIC, precision=ox.FP32_IR,
                        tol=1e-13, keep_steps=True, system=(pat, s))
        assert s.spmv_format(True) == 25, name
        for x, xr in zip(res["steps"], ref["steps"]):
            assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < SOL_TOL
        s.close()
    sc = SCEN["quad"]
    pat, s = W.make_system(sc, ox.FP32_IR)
    res = W.run_gpu(sc, method=ox.BICGSTAB, dirichlet=ox.DIRICHLET_ROW, precision=ox.FP32_IR,
                    system=(pat, s))
    assert all(i["converged"] == 1 for i in res["info"])
    assert s.spmv_format(True) != 25
    s.close()


@pytest.mark.parametrize("name,method,bc", [("quad", ox.CG, ox.DIRICHLET_SYMMETRIC),
                                            ("quad", ox.BICGSTAB, ox.DIRICHLET_ROW),
                                            ("hex", ox.CG, ox.DIRICHLET_SYMMETRIC),
                                            ("tet", ox.BICGSTAB, ox.DIRICHLET_ROW)])
def test_solves_identical_in_every_tuned_format(name, method, bc):
    """Every format the tuner may pick (and the plain CRS format) sums rows
    in CRS order and hands each row to the same thread of the same block,
    so whole solves -- iterates, iteration counts, residuals -- are
    bit-identical whichever format the timing-based tuner chose."""
    sc = SCEN[name]
    ref = None
    for fmt in (9, 3, 5, 7, 11, 12, 13, 14, 17, 18, 21, 22, 19, 20):
        pat, s = W.make_system(sc)
        try:
            s.set_spmv_variant(fmt)
        except ox.OgsxError:
            s.close()
            continue
        res = W.run_gpu(sc, method=method, dirichlet=bc, keep_steps=True, system=(pat, s))
        its = [i["iterations"] for i in res["info"]]
        if ref is None:
            ref = res
        else:
            assert its == [i["iterations"] for i in ref["info"]], fmt
            assert [i["error"] for i in res["info"]] == [i["error"] for i in ref["info"]], fmt
            for x, xr in zip(res["steps"], ref["steps"]):
                assert np.array_equal(x, xr), fmt
        s.close()


def test_solves_identical_across_formats_full_size_tet():
    """The same identity at the full 96^3 x 6 tet heat size (912 673 DOF,
    rows of 15 entries, BiCGStab with row elimination), where the fused
    solver kernels run many blocks per row batch: the plain CRS format and
    every stencil-dictionary format give the same iterates."""
    sc = W.tet_heat(96, n_steps=2)
    pat = ox.Pattern(sc.n, sc.mesh.etype, sc.mesh.conn)
    ref = None
    for fmt in (9, 13, 14, 17, 18, 21, 22):
        s = ox.System(pat, ox.FP64)
        s.set_elements(**sc.element_kwargs())
        s.set_spmv_variant(fmt)
        res = W.run_gpu(sc, method=ox.BICGSTAB, dirichlet=ox.DIRICHLET_ROW, system=(pat, s))
        s.close()
        its = [i["iterations"] for i in res["info"]]
        assert all(i["converged"] == 1 for i in res["info"]), fmt
        if ref is None:
            ref = (its, res["x"])
        else:
            assert its == ref[0], fmt
            assert np.array_equal(res["x"], ref[1]), fmt
    pat.close()


def test_stored_initial_condition_on_device():
    """A job restarted from the stored vector (OGSX_VEC_IC, copied on the
    device, including the transient restart after the steady step) gives
    bit-for-bit the job that uploads u0 from the host, twice in a row."""
    sc = SCEN["hex"]
    assert any(st.reset_to_ic for st in sc.steps)
    pat, s = W.make_system(sc, ox.FP64)
    kw = dict(method=ox.CG, dirichlet=ox.DIRICHLET_SYMMETRIC, precision=ox.FP64, system=(pat, s))
    host = W.run_gpu(sc, **kw)["x"]
    s.set_vector(ox.VEC_IC, sc.u0)
    assert np.array_equal(s.get_vector(ox.VEC_IC), sc.u0)
    for _ in range(2):
        out = W.run_gpu(sc, ic_on_device=True, fetch_x=False, **kw)
        assert out["x"] is None
        assert np.array_equal(s.get_vector(ox.VEC_X), host)
    s.close()


def test_stencil_dictionary_needs_repeated_patterns():
    """A pattern with more than 255 distinct column-offset lists has no
    dictionary: the formats are refused, the others still work."""
    rng = np.random.default_rng(3)
    n = 3000
    chain = np.stack([np.arange(n - 1), np.arange(1, n)], 1)  # every node in an element
    extra = np.stack([np.arange(n), rng.integers(0, n, n)], 1)
    conn = np.concatenate([chain, extra[extra[:, 0] != extra[:, 1]]])
    pat = ox.Pattern(n, ox.LINE2, conn)
    s = ox.System(pat, ox.FP64)
    with pytest.raises(ox.OgsxError):
        s.set_spmv_variant(13)
    with pytest.raises(ox.OgsxError):
        s.set_spmv_variant(19)  # slice stencils far beyond 16 triples
    s.set_spmv_variant(12)
    s.close()


def _tight_reference(sc):
    """Oracle solution chain at tol 1e-14: CG on the symmetric elimination for
    LIQUID (same solution as the reference's row-only system), BiCGStab on
    the row-only system for the nonsymmetric HEAT operator."""
    from oracle import ffi
    if sc.process == ox.HEAT_TRANSPORT:
        ref = driver.run_oracle(sc, ffi, method=2, symmetric=False, tol=1e-14, keep_steps=True)
    else:
        ref = driver.run_oracle(sc, ffi, method=5, symmetric=True, tol=1e-14, keep_steps=True)
    assert all(i["error"] < 1e-13 for i in ref["info"])
    return ref


def test_reference_solution_chain_is_exact():
    """The tight oracle chain equals dense solves of the oracle-assembled
    row-only systems (the reference's own linear systems)."""
    from oracle import ffi
    for name in ["hex", "tet"]:
        sc = SCEN[name]
        ref = _tight_reference(sc)
        row = driver.run_oracle(sc, ffi, method=2, symmetric=False, tol=1e-14, keep_steps=True,
                                max_steps=1, stop_after_assembly=True)
        eq = row["eqs"]
        a, b = row["systems"][0]
        eq.a[:] = a
        x = np.linalg.solve(eq.dense(), b)
        assert np.linalg.norm(ref["steps"][0] - x) / np.linalg.norm(x) < 1e-12


@pytest.mark.parametrize("name", ["quad", "hex", "tet", "hex_aniso_grav", "quad_gw", "hex_ng4",
                                  "quad_ng1", "quad_geo", "quad_heat_geo", "line_geo",
                                  "hex_sink"])
@pytest.mark.parametrize("method,bc,prec,ir_mode", [
    (ox.CG, ox.DIRICHLET_SYMMETRIC, ox.FP64, None),
    (ox.BICGSTAB, ox.DIRICHLET_ROW, ox.FP64, None),
    (ox.CG, ox.DIRICHLET_SYMMETRIC, ox.FP32_IR, ox.IR_RELIABLE),
    (ox.CG, ox.DIRICHLET_SYMMETRIC, ox.FP32_IR, ox.IR_RELIABLE_SR),
    (ox.CG, ox.DIRICHLET_SYMMETRIC, ox.FP32_IR, ox.IR_RESTART),
    (ox.BICGSTAB, ox.DIRICHLET_ROW, ox.FP32_IR, None),
])
def test_solution_parity(name, method, bc, prec, ir_mode):
    from oracle import ffi
    sc = SCEN[name]
    if method == ox.CG and sc.process == ox.HEAT_TRANSPORT:
        pytest.skip("advection makes A nonsymmetric: CG is not applicable")
    if method == ox.BICGSTAB:
        # The reference BiCGStab breaks down exactly ((r0,r) = 0 at iteration
        # 2) when b lives only on row-eliminated Dirichlet rows and x0 = 0.
        # There the GPU must report the same breakdown as the oracle.
        # The breakdown hinges on exact cancellations, so whether the GPU's
        # FMA-rounded recurrence breaks down at the same step is not defined
        # by the reference: such systems are left out of the parity check.
        same = driver.run_oracle(sc, ffi, method=2, symmetric=False, tol=1e-13)
        broke = [k for k, i in enumerate(same["info"]) if i["ret"] == 0 and i["error"] > 1e-13]
        if broke:
            pytest.skip(f"reference BiCGStab breaks down at step {broke[0]} of this system")
    ref = _tight_reference(sc)
    res = W.run_gpu(sc, method=method, dirichlet=bc, precision=prec, tol=1e-13, keep_steps=True,
                    ir_mode=ir_mode)
    print(f"[solve] {name} m={method} bc={bc} prec={prec} ir={ir_mode} "
          f"iters={[i['iterations'] for i in res['info']]}", flush=True)
    for k, (x, xr) in enumerate(zip(res["steps"], ref["steps"])):
        rel = np.linalg.norm(x - xr) / np.linalg.norm(xr)
        assert rel < SOL_TOL, (k, rel)
    assert all(i["converged"] in (1, 2) for i in res["info"])


def test_same_algorithm_iteration_counts():
    """GPU CG and BiCGStab follow the reference recurrences: iteration counts
    match the oracle's on the same system up to rounding-order effects."""
    from oracle import ffi
    sc = SCEN["quad"]
    for method, sym in [(ox.CG, True), (ox.BICGSTAB, False)]:
        ref = driver.run_oracle(sc, ffi, method=method, symmetric=sym)
        res = W.run_gpu(sc, method=method,
                        dirichlet=ox.DIRICHLET_SYMMETRIC if sym else ox.DIRICHLET_ROW)
        for a, b in zip(res["info"], ref["info"]):
            assert abs(a["iterations"] - b["iterations"]) <= max(2, 0.03 * b["iterations"])
            assert a["ret"] == b["ret"] or method == ox.CG


def test_bmskel_golden_on_gpu():
    """The reference's bmskel deck: T(POINT0) = 274 after one step."""
    sc = SCEN["bmskel"]
    res = W.run_gpu(sc, method=ox.BICGSTAB, dirichlet=ox.DIRICHLET_ROW)
    assert res["x"][0] == 274.0
    # node 1 after the Crank-Nicolson step, from the 2x2 system in closed form
    from closed_forms import bmskel_node1
    assert abs(res["x"][1] - bmskel_node1()) <= 1e-12 * abs(bmskel_node1())


def test_write_bin_matches_oracle(tmp_path):
    from oracle import ffi
    sc = SCEN["quad"]
    A, b, Ar, br, ptr, s, eq = _assemble_both(sc, ox.DIRICHLET_ROW)
    s.write_bin(tmp_path / "gpu.bin")
    eq.write_bin(tmp_path / "ref.bin")
    n1, p1, c1, a1, b1 = ffi.read_bin(tmp_path / "gpu.bin")
    n2, p2, c2, a2, b2 = ffi.read_bin(tmp_path / "ref.bin")
    assert n1 == n2 and np.array_equal(p1, p2) and np.array_equal(c1, c2)
    assert _row_scaled_close(a1, a2, p2, ENTRY_TOL) < ENTRY_TOL
    assert np.abs(b1 - b2).max() <= ENTRY_TOL * np.abs(b2).max()


def test_deterministic_and_check_norm_rhs():
    sc = SCEN["hex"]
    r1 = W.run_gpu(sc, method=ox.CG)
    r2 = W.run_gpu(sc, method=ox.CG)
    assert np.array_equal(r1["x"], r2["x"])
    pat, s = r1["system"]
    # a right-hand side 1e-12 of the last one is skipped (CheckNormRHS)
    b = s.get_vector(ox.VEC_B)
    x0 = s.get_vector(ox.VEC_X)
    s.set_vector(ox.VEC_B, b * 1e-12)
    info = s.solve(method=ox.CG, tol=1e-10)
    assert info["converged"] == 2 and np.array_equal(s.get_vector(ox.VEC_X), x0)


def test_max_iter_return_values():
    sc = SCEN["quad"]
    for method, ret in [(ox.CG, 0), (ox.BICGSTAB, 4)]:
        pat, s = W.make_system(sc)
        s.set_vector(ox.VEC_UOLD, sc.u0)
        s.initialize()
        st = sc.steps[0]
        s.assemble(st.dt, st.theta)
        s.add_rhs(sc.st_idx, sc.st_val)
        s.set_dirichlet(sc.bc_idx, sc.bc_val, ox.DIRICHLET_SYMMETRIC)
        info = s.solve(method=method, tol=1e-10, max_iter=3)
        assert info["converged"] == 0 and info["ret"] == ret


def test_full_size_quad_config_vs_oracle():
    """The CPU-runnable configuration at full size (200 x 200 quads, 10
    steps): every step's solution against the oracle's BiCGStab."""
    from oracle import ffi
    sc = W.quad_flow()
    ref = driver.run_oracle(sc, ffi, method=2, symmetric=False, tol=1e-14, keep_steps=True)
    assert all(i["error"] < 1e-13 for i in ref["info"])
    res = W.run_gpu(sc, method=ox.CG, tol=1e-13, keep_steps=True)
    for x, xr in zip(res["steps"], ref["steps"]):
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < SOL_TOL


def test_full_size_tet_heat_properties():
    """Heat transport on 96^3 cubes x 6 tets (912,673 nodes, 13.5M nonzeros,
    divergence-free Gauss-point velocity, grid Peclet 1-20), first
    Crank-Nicolson step: size-independent properties. CRS nnz closed form;
    the steady operator (conduction + advection with a divergence-free
    velocity) annihilates constants; every SpMV format the pattern admits
    gives the same bits; FP64 BiCGStab on the reference's row-only system and
    the FP32-inside-FP64 refinement reach true FP64 residual <= 1e-10 and
    agree."""
    import scipy.sparse as sp
    n = 96
    sc = W.tet_heat(n, n_steps=1)
    pat, s = W.make_system(sc)
    q = n + 1
    assert pat.nnz == q ** 3 + 2 * (3 * q * q * n + 3 * q * n * n + n ** 3)
    ptr, col, _ = pat.export()
    st = sc.steps[0]
    s.set_vector(ox.VEC_UOLD, sc.u0)
    s.set_vector(ox.VEC_X, sc.u0)
    s.initialize()
    s.assemble(st.dt, st.theta, True)
    A0 = s.get_matrix()
    csr = sp.csr_matrix((A0, col, ptr), shape=(sc.n, sc.n))
    assert np.abs(csr @ np.ones(sc.n)).max() <= 1e-12 * np.abs(A0).max() * 15
    x = np.random.default_rng(96).standard_normal(sc.n)
    s.set_spmv_variant(5)
    y5 = s.spmv(x)
    tried = []
    for f in (3, 13, 14, 17, 18, 21, 22, 19, 20):
        try:
            s.set_spmv_variant(f)
        except ox.OgsxError:
            continue
        tried.append(f)
        assert np.array_equal(s.spmv(x), y5), f
    print(f"[full tet] formats compared with 5: {tried}", flush=True)
    s.set_spmv_variant(0)
    s.initialize()
    s.assemble(st.dt, st.theta)
    s.set_dirichlet(sc.bc_idx, sc.bc_val, ox.DIRICHLET_ROW)
    A = s.get_matrix()
    b = s.get_vector(ox.VEC_B)
    csr = sp.csr_matrix((A, col, ptr), shape=(sc.n, sc.n))
    info = s.solve(method=ox.BICGSTAB, tol=1e-10, max_iter=10000)
    x64 = s.get_vector(ox.VEC_X)
    r = np.linalg.norm(b - csr @ x64) / np.linalg.norm(b)
    print(f"[full tet] iters={info['iterations']} relres={r:.3e}", flush=True)
    assert info["converged"] == 1 and r <= 1e-10 * 1.5
    pat2, s2 = W.make_system(sc, ox.FP32_IR)
    s2.set_vector(ox.VEC_UOLD, sc.u0)
    s2.set_vector(ox.VEC_X, sc.u0)
    s2.initialize()
    s2.assemble(st.dt, st.theta)
    s2.set_dirichlet(sc.bc_idx, sc.bc_val, ox.DIRICHLET_ROW)
    info2 = s2.solve(method=ox.BICGSTAB, tol=1e-10, max_iter=10000)
    x32 = s2.get_vector(ox.VEC_X)
    r2 = np.linalg.norm(b - csr @ x32) / np.linalg.norm(b)
    print(f"[full tet IR] iters={info2['iterations']} outer={info2['outer']} relres={r2:.3e}",
          flush=True)
    assert info2["converged"] == 1 and r2 <= 1e-10
    # error of both solutions against a tight FP64 solve of the same system
    # (the device FP64 BiCGStab, oracle-pinned at small sizes), continued
    # from x64 to a 1e-13 residual
    info3 = s.solve(method=ox.BICGSTAB, tol=1e-13, max_iter=10000)
    xr = s.get_vector(ox.VEC_X)
    r3 = np.linalg.norm(b - csr @ xr) / np.linalg.norm(b)
    e64 = np.linalg.norm(x64 - xr) / np.linalg.norm(xr)
    e32 = np.linalg.norm(x32 - xr) / np.linalg.norm(xr)
    print(f"[full tet] tight ref relres={r3:.3e} err fp64={e64:.3e} err ir={e32:.3e}", flush=True)
    assert r3 <= 1e-12
    assert e64 <= TET_ERR_BOUND and e32 <= TET_ERR_BOUND
    s.close()
    s2.close()


def test_full_size_hex_properties():
    """128^3 hexes (2.1M nodes, 57M nonzeros): size-independent properties.
    CRS nnz closed form; symmetric elimination leaves A symmetric; the steady
    operator annihilates constants on interior rows; the FP64 CG and the FP32
    iterative-refinement solutions have true FP64 residual <= 1e-10 and agree."""
    import scipy.sparse as sp
    n = 128
    sc = W.hex_flow(n, n_transient=0)
    pat, s = W.make_system(sc)
    assert pat.nnz == (3 * (n + 1) - 2) ** 3
    ptr, col, _ = pat.export()
    st = sc.steps[0]
    s.set_vector(ox.VEC_UOLD, sc.u0)
    s.set_vector(ox.VEC_X, sc.u0)
    s.initialize()
    s.assemble(st.dt, st.theta, True)
    A0 = s.get_matrix()
    csr = sp.csr_matrix((A0, col, ptr), shape=(sc.n, sc.n))
    rs = np.abs(csr @ np.ones(sc.n))
    assert rs.max() <= 1e-12 * np.abs(A0).max() * 27
    s.set_dirichlet(sc.bc_idx, sc.bc_val, ox.DIRICHLET_SYMMETRIC)
    A = s.get_matrix()
    b = s.get_vector(ox.VEC_B)
    csr = sp.csr_matrix((A, col, ptr), shape=(sc.n, sc.n))
    assert abs(csr - csr.T).max() == 0.0
    info = s.solve(method=ox.CG, tol=1e-10, max_iter=100000)
    x64 = s.get_vector(ox.VEC_X)
    r = np.linalg.norm(b - csr @ x64) / np.linalg.norm(b)
    print(f"[full hex] iters={info['iterations']} relres={r:.3e}", flush=True)
    assert info["converged"] == 1 and r <= 1e-10 * 1.5
    pat2, s2 = W.make_system(sc, ox.FP32_IR)
    s2.set_vector(ox.VEC_UOLD, sc.u0)
    s2.set_vector(ox.VEC_X, sc.u0)
    s2.initialize()
    s2.assemble(st.dt, st.theta, True)
    s2.set_dirichlet(sc.bc_idx, sc.bc_val, ox.DIRICHLET_SYMMETRIC)
    info2 = s2.solve(method=ox.CG, tol=1e-10, max_iter=100000)
    x32 = s2.get_vector(ox.VEC_X)
    r2 = np.linalg.norm(b - csr @ x32) / np.linalg.norm(b)
    print(f"[full hex IR] iters={info2['iterations']} outer={info2['outer']} relres={r2:.3e}",
          flush=True)
    assert info2["converged"] == 1 and r2 <= 1e-10
    # error of both solutions against a tight FP64 solve of the same system
    # (the device FP64 CG, oracle-pinned at small sizes), continued from x64
    # to a 1e-13 residual
    info3 = s.solve(method=ox.CG, tol=1e-13, max_iter=100000)
    xr = s.get_vector(ox.VEC_X)
    r3 = np.linalg.norm(b - csr @ xr) / np.linalg.norm(b)
    e64 = np.linalg.norm(x64 - xr) / np.linalg.norm(xr)
    e32 = np.linalg.norm(x32 - xr) / np.linalg.norm(xr)
    print(f"[full hex] tight ref iters={info3['iterations']} relres={r3:.3e} "
          f"err fp64={e64:.3e} err ir={e32:.3e}", flush=True)
    assert r3 <= 1e-12
    assert e64 <= HEX_ERR_BOUND and e32 <= HEX_ERR_BOUND
