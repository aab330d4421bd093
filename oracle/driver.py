"""Runs a scenario through the oracle in the order CRFProcess::Execute calls
the path (FEM/rf_pcs.cpp:4695-4853).  Test infrastructure only."""
from __future__ import annotations

import numpy as np


def run_oracle(sc, ffi, *, method: int = 2, symmetric: bool = False, keep_steps: bool = False,
               tol=None, max_steps=None, stop_after_assembly=False):
    """method 2 = BiCGStab (the reference's own solve for row-only Dirichlet),
    5 = CG (needs symmetric=True).  Returns x, per-step info and, when asked,
    each step's (A, b) before the solve."""
    ptr, col, diag = ffi.pattern(sc.n, sc.mesh.conn)
    eq = ffi.OracleEqs(ptr, col, diag)
    eq.config(method, 1, sc.tol if tol is None else tol, sc.max_iter)
    prob = ffi.OracleProblem(sc.mesh, process=sc.process, mass_coef=sc.mass_coef, cond=sc.cond,
                             cond_stride=sc.cond_stride, n_gauss=sc.n_gauss, tuf=sc.tuf, mu=sc.mu,
                             geo_area=sc.geo_area, adv_coef=sc.adv_coef, vel=sc.vel,
                             rho_g=sc.rho_g, u_old=sc.u0)
    u_old = sc.u0.copy()
    eq.x[:] = sc.u0
    infos, systems, xs = [], [], []
    steps = sc.steps if max_steps is None else sc.steps[:max_steps]
    for st in steps:
        if st.reset_to_ic:
            u_old = sc.u0.copy()
            eq.x[:] = sc.u0
        prob.set(dt=st.dt, theta=st.theta, steady=st.steady, u_old=u_old)
        eq.initialize()
        eq.assemble(prob)
        if sc.st_idx.size:
            eq.add_rhs(sc.st_idx, sc.st_val)
        eq.set_known_x(sc.bc_idx, sc.bc_val, symmetric=symmetric)
        if keep_steps:
            systems.append((eq.a.copy(), eq.b.copy()))
        if stop_after_assembly:
            break
        infos.append(eq.solve())
        u_old = eq.x.copy()
        if keep_steps:
            xs.append(eq.x.copy())
    return dict(x=eq.x.copy(), info=infos, ptr=ptr, col=col, diag=diag, systems=systems,
                steps=xs, eqs=eq)


def run_oracle_coupled(flow, heat, ffi, *, tol=None):
    """LIQUID_FLOW then HEAT_TRANSPORT in each time step, the heat step
    advected by the Gauss-point velocity of the new pressure (Cal_Velocity
    via CalIntegrationPointValue, FEM/rf_pcs.cpp:5915, called after the
    flow solve).  Flow: CG with symmetric Dirichlet; heat (non-symmetric
    with advection): BiCGStab with row-only Dirichlet.  Returns both final
    solutions, the per-step infos and the last velocity."""
    def setup(sc, method):
        ptr, col, diag = ffi.pattern(sc.n, sc.mesh.conn)
        eq = ffi.OracleEqs(ptr, col, diag)
        eq.config(method, 1, sc.tol if tol is None else tol, sc.max_iter)
        prob = ffi.OracleProblem(sc.mesh, process=sc.process, mass_coef=sc.mass_coef,
                                 cond=sc.cond, cond_stride=sc.cond_stride, n_gauss=sc.n_gauss,
                                 tuf=sc.tuf, mu=sc.mu, geo_area=sc.geo_area,
                                 adv_coef=sc.adv_coef, vel=sc.vel, rho_g=sc.rho_g, u_old=sc.u0)
        eq.x[:] = sc.u0
        return eq, prob

    def step(sc, eq, prob, st, u_old, symmetric):
        prob.set(dt=st.dt, theta=st.theta, steady=st.steady, u_old=u_old)
        eq.initialize()
        eq.assemble(prob)
        if sc.st_idx.size:
            eq.add_rhs(sc.st_idx, sc.st_val)
        eq.set_known_x(sc.bc_idx, sc.bc_val, symmetric=symmetric)
        info = eq.solve()
        return info, eq.x.copy()

    eqf, pf = setup(flow, 5)
    eqh, ph = setup(heat, 2)
    uf, uh = flow.u0.copy(), heat.u0.copy()
    infos_f, infos_h, vel = [], [], None
    for k in range(len(flow.steps)):
        info, uf = step(flow, eqf, pf, flow.steps[k], uf, True)
        infos_f.append(info)
        vel = pf.gauss_velocity(uf)
        ph.set(vel=vel)
        info, uh = step(heat, eqh, ph, heat.steps[k], uh, False)
        infos_h.append(info)
    return dict(x_flow=uf, x_heat=uh, info_flow=infos_f, info_heat=infos_h, vel=vel)
