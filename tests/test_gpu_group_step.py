"""Batched time steps of a group (ogsx_group_step) against the per-system
calls of CRFProcess::Execute (FEM/rf_pcs.cpp:4695-4853: copy-back,
Initialize, element loop, source terms, Dirichlet): after every step each
member's A, b, solution and solve info must be bit-identical, for every
element type, both processes, full tensors, gravity, advection, steady
steps, restarts from the initial condition, both Dirichlet modes and both
precisions, in one group of mixed members."""
import numpy as np
import pytest

import ogs5_amd as ox
from ogs5_amd import _lib as L
from ogs5_amd import workloads as W

import test_gpu_parity as P

pytestmark = pytest.mark.gpu


def _mix():
    scs = P._scenarios()
    names = ["quad", "hex", "tet", "hex_aniso_grav", "quad_gw", "hex_sink", "quad_geo",
             "quad_heat_geo", "line_geo", "quad_ng4"]
    out = [scs[k] for k in names]
    out.append(W.quad_flow(33, 21, seed=71, sigma_ln_k=2.0, n_steps=4))
    return out


def _drive(scs, prec, dirichlet, method, batched, **cfg):
    systems = [W.make_system(sc, prec) for sc in scs]
    g = L.Group(prec)
    for sc, (_, s) in zip(scs, systems):
        g.add(s)
        s.set_vector(L.VEC_UOLD, sc.u0)
        s.set_vector(L.VEC_X, sc.u0)
    rec = [[] for _ in scs]
    for k in range(max(len(sc.steps) for sc in scs)):
        live = [i for i, sc in enumerate(scs) if k < len(sc.steps)]
        W._start_steps(g, scs, systems, live, [k] * len(scs), dirichlet, batched, method=method,
                       tol=1e-10, **cfg)
        while g.advance() > 0:
            pass
        for i in live:
            s = systems[i][1]
            active, info = L.solve_status(s)
            assert not active
            rec[i].append((s.get_matrix(), s.get_vector(L.VEC_B), s.get_vector(L.VEC_X), info))
    g.close()
    for _, s in systems:
        s.close()
    return rec


CASES = [
    (ox.FP64, ox.DIRICHLET_SYMMETRIC, ox.BICGSTAB, {}),
    (ox.FP64, ox.DIRICHLET_ROW, ox.BICGSTAB, {}),
    (ox.FP32_IR, ox.DIRICHLET_SYMMETRIC, ox.BICGSTAB, {}),
]


@pytest.mark.parametrize("prec,bc,method,cfg", CASES)
def test_group_step_equals_per_system_calls(prec, bc, method, cfg):
    scs = _mix()
    one = _drive(scs, prec, bc, method, False, **cfg)
    many = _drive(scs, prec, bc, method, True, **cfg)
    batched = 0
    for i, sc in enumerate(scs):
        for k, (a, b) in enumerate(zip(one[i], many[i])):
            batched += W.batchable(sc, k)
            for name, u, v in (("A", a[0], b[0]), ("b", a[1], b[1]), ("x", a[2], b[2])):
                assert np.array_equal(u, v), f"{sc.name} step {k}: {name} differs"
            assert a[3]["iterations"] == b[3]["iterations"], (sc.name, k)
            assert a[3]["error"] == b[3]["error"], (sc.name, k)
            assert a[3]["converged"] == b[3]["converged"], (sc.name, k)
    print(f"[group_step] prec={prec} bc={bc} m={method}: {batched} batched member steps "
          f"bit-identical", flush=True)
    assert batched >= 15


def test_group_step_cg_on_symmetric_members():
    """CG members (symmetric tensors, no advection) in FP32-IR with the
    single-reduction recurrence: the batched step restores the symmetry
    flag the recurrence's choice depends on."""
    scs = [W.quad_flow(25, 19, seed=72, n_steps=3), W.hex_flow(7, seed=73, n_transient=2),
           W.quad_flow(12, 40, seed=74, n_steps=4)]
    one = _drive(scs, ox.FP32_IR, ox.DIRICHLET_SYMMETRIC, ox.CG, False,
                 ir_mode=ox.IR_RELIABLE_SR)
    many = _drive(scs, ox.FP32_IR, ox.DIRICHLET_SYMMETRIC, ox.CG, True,
                  ir_mode=ox.IR_RELIABLE_SR)
    for i in range(len(scs)):
        for a, b in zip(one[i], many[i]):
            assert np.array_equal(a[2], b[2])
            assert a[3]["iterations"] == b[3]["iterations"]


def test_group_step_continuous_run_matches_lock_step():
    """The continuous runner with batched steps (members start their next
    step as soon as their solve ends) against lock-step per-system steps."""
    scs = [W.quad_flow(n, n + 3, seed=80 + n, sigma_ln_k=1.5, n_steps=3) for n in (9, 26, 41)]
    ref = W.run_group(scs, precision=ox.FP32_IR, keep_steps=True)
    got = W.run_group(scs, precision=ox.FP32_IR, keep_steps=True, continuous=True,
                      batched_steps=True, batch_iters=8)
    for k in range(len(scs)):
        for xa, xb in zip(ref["steps"][k], got["steps"][k]):
            assert np.array_equal(xa, xb)


def test_group_step_rejects_long_rows_and_strangers():
    m = P._star_mesh(40)
    E = m.n_elem
    star = W.Scenario("star", m, ox.LIQUID_FLOW, np.full(E, 1e-4), np.full(E, 1e-5), 1,
                      np.array([m.n_nodes - 1], np.int64), np.array([1.0]),
                      np.zeros(0, np.int64), np.zeros(0), np.zeros(m.n_nodes),
                      [W.Step(10.0, 1.0)] * 2)
    quad = W.quad_flow(8, 6, seed=90, n_steps=2)
    (_, s_star), (_, s_quad) = W.make_system(star), W.make_system(quad)
    (_, s_out) = W.make_system(W.quad_flow(5, 5, seed=91, n_steps=1))
    g = L.Group(ox.FP64)
    g.add(s_star)
    g.add(s_quad)
    p = dict(dt=10.0, theta=1.0, dirichlet=ox.DIRICHLET_SYMMETRIC)
    for s in (s_star, s_quad):
        s.initialize()
        s.assemble(10.0, 1.0, False)  # two-pass buffers exist where the rows allow
    with pytest.raises(L.OgsxError, match="two-pass"):
        g.step([s_star], [p], method=ox.CG, tol=1e-10)
    with pytest.raises(L.OgsxError, match="not a member"):
        g.step([s_out], [p], method=ox.CG, tol=1e-10)
    with pytest.raises(L.OgsxError, match="listed twice"):
        g.step([s_quad, s_quad], [p, p], method=ox.CG, tol=1e-10)
    with pytest.raises(L.OgsxError, match="dt must be positive"):
        g.step([s_quad], [dict(p, dt=0.0)], method=ox.CG, tol=1e-10)
    # nothing was started: the quad member steps normally afterwards (from
    # u_old = 1, so that b != 0)
    s_quad.set_vector(L.VEC_X, np.ones(quad.n))
    g.step([s_quad], [dict(p, shift=True)], method=ox.CG, tol=1e-10)
    while g.advance() > 0:
        pass
    assert L.solve_status(s_quad)[1]["converged"] == 1
    g.close()
    for s in (s_star, s_quad, s_out):
        s.close()
