"""Synthetic scenarios of the benchmark configurations and the time-step
driver that calls the C ABI in the order CRFProcess::Execute does
(FEM/rf_pcs.cpp:4695-4853): Initialize -> x0 = u(t_n) -> GlobalAssembly ->
IncorporateSourceTerms -> IncorporateBoundaryConditions -> Solver -> copy
back u(t_n+1) = x.

Scenarios (all seeded, all synthetic):
  quad_flow   2-D transient confined groundwater flow on bilinear quads
              (200 x 200 over a 1000 m square is the CPU-runnable config)
  hex_flow    3-D liquid flow on trilinear hexes (128^3 is the headline
              config): one steady solve, then transient steps from the IC
  tet_heat    advection-diffusion heat transport on 6-tet-split cubes
  bmskel      the reference's own 1-element HEAT_TRANSPORT deck
              (tests/data/bmskel)
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib as L
from .fields import divergence_free_velocity, gauss_point_coords, lognormal_conductivity
from .mesh import Mesh, hex_mesh, line_mesh, quad_mesh, tet_mesh


@dataclass
class Step:
    dt: float
    theta: float
    steady: bool = False
    reset_to_ic: bool = False  # start from the initial condition again


@dataclass
class Scenario:
    name: str
    mesh: Mesh
    process: int
    mass_coef: np.ndarray
    cond: np.ndarray
    cond_stride: int
    bc_idx: np.ndarray
    bc_val: np.ndarray
    st_idx: np.ndarray
    st_val: np.ndarray
    u0: np.ndarray
    steps: List[Step]
    tol: float = 1e-10
    max_iter: int = 100000
    n_gauss: int = 3
    tuf: float = 1.0
    mu: float = 1.0
    geo_area: float = 1.0
    adv_coef: float = 0.0
    vel: Optional[np.ndarray] = None
    rho_g: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.mesh.n_nodes

    def element_kwargs(self) -> dict:
        m = self.mesh
        return dict(process=self.process, dim=m.dim, x=m.x, y=m.y if m.dim >= 2 else None,
                    z=m.z if m.dim >= 3 else None, mass_coef=self.mass_coef, cond=self.cond,
                    cond_stride=self.cond_stride, n_gauss=self.n_gauss, tuf=self.tuf, mu=self.mu,
                    geo_area=self.geo_area, adv_coef=self.adv_coef, vel=self.vel,
                    rho_g=self.rho_g)


def quad_flow(nx: int = 200, ny: int = 200, *, length: float = 1000.0, seed: int = 1,
              mean_ln_k: float = np.log(1e-4), sigma_ln_k: float = 1.0, corr_len: float = 50.0,
              storage: float = 1e-4, head_left: float = 10.0, head_right: float = 0.0,
              sink: float = -0.01, dt: float = 3600.0, n_steps: int = 10) -> Scenario:
    """2-D transient confined flow (the 200x200 / 1000 m configuration by default).

    Dirichlet head on the left and right edges, one point sink at the centre
    node, implicit Euler (theta = 1), x0 = previous time level.
    """
    m = quad_mesh(nx, ny, length, length * ny / nx)
    K = lognormal_conductivity(m, mean_ln_k, sigma_ln_k, corr_len, seed)
    left = m.nodes_where(0, 0.0)
    right = m.nodes_where(0, length)
    bc_idx = np.concatenate([left, right])
    bc_val = np.concatenate([np.full(left.size, head_left), np.full(right.size, head_right)])
    centre = m.node_index(nx // 2, ny // 2)
    return Scenario(f"quad_flow_{nx}x{ny}", m, L.LIQUID_FLOW, np.full(m.n_elem, storage), K, 1,
                    bc_idx.astype(np.int64), bc_val, np.array([centre], np.int64),
                    np.array([sink]), np.zeros(m.n_nodes), [Step(dt, 1.0)] * n_steps,
                    meta=dict(seed=seed, sigma_ln_k=sigma_ln_k))


def hex_flow(n: int = 128, *, seed: int = 2, mean_ln_k: float = np.log(1e-4),
             sigma_ln_k: float = 2.0, corr_cells: float = 8.0, storage: float = 1e-4,
             dh: float = 1.0, dt: float = 100.0, n_transient: int = 5,
             steady: bool = True) -> Scenario:
    """3-D liquid flow on n^3 unit hexes: head dh on x = 0, 0 on x = n.

    Steps: one steady solve (zero storage) from h = 0, then n_transient
    implicit-Euler steps that restart from the initial condition h = 0.
    """
    m = hex_mesh(n, n, n)
    K = lognormal_conductivity(m, mean_ln_k, sigma_ln_k, corr_cells, seed)
    left = m.nodes_where(0, 0.0)
    right = m.nodes_where(0, float(n))
    bc_idx = np.concatenate([left, right]).astype(np.int64)
    bc_val = np.concatenate([np.full(left.size, dh), np.zeros(right.size)])
    steps = []
    if steady:
        steps.append(Step(dt, 1.0, steady=True))
    for k in range(n_transient):
        steps.append(Step(dt, 1.0, reset_to_ic=(k == 0 and steady)))
    return Scenario(f"hex_flow_{n}^3", m, L.LIQUID_FLOW, np.full(m.n_elem, storage), K, 1,
                    bc_idx, bc_val, np.zeros(0, np.int64), np.zeros(0), np.zeros(m.n_nodes),
                    steps, meta=dict(seed=seed, sigma_ln_k=sigma_ln_k))


def tet_heat(n: int = 96, *, seed: int = 3, conductivity: float = 2.0, rho_c: float = 2.5e6,
             fluid_rho_c: float = 4.2e6, peclet_range=(1.0, 20.0), t0: float = 10.0,
             t_inlet: float = 60.0, dt: float = 1.0e5, n_steps: int = 20,
             theta: float = 0.5) -> Scenario:
    """Advection-diffusion of heat on n^3 cubes split into 6 tets each.

    Gauss-point velocity v = curl(0,0,psi) (divergence-free), scaled so the
    grid Peclet c_f rho_f |v| h / (2 lambda) spans peclet_range; inlet face
    x = 0 held at t_inlet; Crank-Nicolson (theta = 0.5).
    """
    m = tet_mesh(n, n, n)
    h = m.h[0]
    pe_lo, pe_hi = peclet_range
    # |v| in [u(1-a), u(1+a)] with (1+a)/(1-a) = pe_hi/pe_lo
    a = (pe_hi / pe_lo - 1.0) / (pe_hi / pe_lo + 1.0)
    u_mean = pe_lo * 2.0 * conductivity / (fluid_rho_c * h) / (1.0 - a)
    gp = gauss_point_coords(m, 3)  # [E, 5, 3]
    v = divergence_free_velocity(gp, u_mean, a * 0.999, float(n) * h, seed)  # [E, 5, 3]
    vel = np.ascontiguousarray(np.transpose(v, (0, 2, 1)))  # [E][3][ngp]
    inlet = m.nodes_where(0, 0.0)
    speed = np.linalg.norm(v, axis=-1)
    pe = fluid_rho_c * speed * h / (2.0 * conductivity)
    return Scenario(f"tet_heat_{n}^3x6", m, L.HEAT_TRANSPORT, np.full(m.n_elem, rho_c),
                    np.full(m.n_elem, conductivity), 1, inlet, np.full(inlet.size, t_inlet),
                    np.zeros(0, np.int64), np.zeros(0), np.full(m.n_nodes, t0),
                    [Step(dt, theta)] * n_steps, adv_coef=fluid_rho_c, vel=vel,
                    meta=dict(seed=seed, peclet_min=float(pe.min()), peclet_max=float(pe.max())))


def coupled_flow_heat(n: int = 32, *, seed: int = 5, mean_ln_k: float = np.log(1e-12),
                      sigma_ln_k: float = 1.0, corr_cells: float = 6.0, mu: float = 1.0e-3,
                      storage: float = 1.0e-7, rho_g: float = 1000.0 * 9.81, dp: float = 2.0e5,
                      conductivity: float = 2.0, rho_c: float = 2.5e6,
                      fluid_rho_c: float = 4.2e6, t0: float = 10.0, t_inlet: float = 60.0,
                      dt: float = 2.0e4, n_steps: int = 10, theta: float = 0.5):
    """Pressure-driven flow through a heterogeneous block that carries heat:
    n^3 cubes split into 6 tets each; LIQUID_FLOW in pressure form with
    gravity (K/mu, rho*g on z) and hydrostatic faces p = dp - rho_g z at
    x = 0 and p = -rho_g z at x = L, starting from p = 0 (implicit Euler);
    HEAT_TRANSPORT advected by the Gauss-point Darcy velocity of each step's
    new pressure, inlet face x = 0 held at t_inlet (Crank-Nicolson).
    Returns (flow, heat); heat.vel is None, the velocity comes from flow."""
    m = tet_mesh(n, n, n)
    lx = float(m.x.max())
    K = lognormal_conductivity(m, mean_ln_k, sigma_ln_k, corr_cells, seed)
    left = m.nodes_where(0, 0.0)
    right = m.nodes_where(0, lx)
    bc_idx = np.concatenate([left, right]).astype(np.int64)
    bc_val = np.concatenate([dp - rho_g * m.z[left], -rho_g * m.z[right]])
    flow = Scenario(f"coupled_flow_{n}^3x6", m, L.LIQUID_FLOW, np.full(m.n_elem, storage), K, 1,
                    bc_idx, bc_val, np.zeros(0, np.int64), np.zeros(0), np.zeros(m.n_nodes),
                    [Step(dt, 1.0)] * n_steps, mu=mu, rho_g=rho_g,
                    meta=dict(seed=seed, sigma_ln_k=sigma_ln_k))
    heat = Scenario(f"coupled_heat_{n}^3x6", m, L.HEAT_TRANSPORT, np.full(m.n_elem, rho_c),
                    np.full(m.n_elem, conductivity), 1, left.astype(np.int64),
                    np.full(left.size, t_inlet), np.zeros(0, np.int64), np.zeros(0),
                    np.full(m.n_nodes, t0), [Step(dt, theta)] * n_steps,
                    adv_coef=fluid_rho_c)
    return flow, heat


def ensemble(count: int = 256, *, side_range=(64, 512), sigma_range=(0.5, 2.5),
             n_steps: int = 10, seed: int = 0) -> List[Scenario]:
    """Independent 2-D flow instances with the quad_flow set-up (1000 m
    square, Dirichlet heads on the x edges, centre sink, 10 implicit steps):
    side s = round(exp(U[ln lo, ln hi])) cells, sigma_lnK ~ U[sigma_range],
    instance i seeded with i."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lo, hi = side_range
    sides = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), count))).astype(int)
    sigmas = rng.uniform(sigma_range[0], sigma_range[1], count)
    return [quad_flow(int(s), int(s), seed=i, sigma_ln_k=float(sg), n_steps=n_steps)
            for i, (s, sg) in enumerate(zip(sides, sigmas))]


@dataclass
class RequestSpec:
    """One served solve request: a flow scenario drawn from the request mix."""
    rid: int
    kind: str        # "quad" (2-D) or "hex" (3-D)
    side: int        # cells per axis
    steps: int       # implicit time steps
    seed: int
    sigma_ln_k: float
    arrival: float   # seconds after the stream starts


def request_stream(count: int, rate: float = 200.0, *, seed: int = 0,
                   quad_sides=(32, 512), hex_sides=(16, 96), steps=(1, 20),
                   sigma_range=(0.5, 2.5)) -> List[RequestSpec]:
    """Poisson arrivals at `rate` per second; each request is a 2-D quad or a
    3-D hex flow problem (even odds) with log-uniform side length, its own
    seeded lognormal K, 1-20 steps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for i in range(count):
        kind = "quad" if rng.random() < 0.5 else "hex"
        lo, hi = quad_sides if kind == "quad" else hex_sides
        side = int(np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        out.append(RequestSpec(i, kind, side, int(rng.integers(steps[0], steps[1] + 1)),
                               1000 + i, float(rng.uniform(*sigma_range)), float(t[i])))
    return out


def request_mesh(kind: str, side: int) -> Mesh:
    """The structured mesh of a request (depends on its kind and size only)."""
    return quad_mesh(side, side, 1000.0, 1000.0) if kind == "quad" else hex_mesh(side, side, side)


def request_scenario(spec: RequestSpec, mesh: Optional[Mesh] = None) -> Scenario:
    """The flow problem of a request: Dirichlet heads on a random pair of
    opposite faces (random values), one sink at a random interior node,
    implicit Euler from h = 0 (2-D: 1000 m square, 50 m correlation, dt
    3600 s; 3-D: unit cells, 8-cell correlation, dt 100 s).  `mesh` may be a
    shared request_mesh(kind, side)."""
    rng = np.random.Generator(np.random.PCG64(spec.seed))
    n = spec.side
    if spec.kind == "quad":
        m = mesh if mesh is not None else request_mesh("quad", n)
        K = lognormal_conductivity(m, np.log(1e-4), spec.sigma_ln_k, 50.0, spec.seed)
        length, dt, q = 1000.0, 3600.0, -0.01
        axes = 2
    else:
        m = mesh if mesh is not None else request_mesh("hex", n)
        K = lognormal_conductivity(m, np.log(1e-4), spec.sigma_ln_k, 8.0, spec.seed)
        length, dt, q = float(n), 100.0, -1e-4
        axes = 3
    axis = int(rng.integers(0, axes))
    lo_face = m.nodes_where(axis, 0.0)
    hi_face = m.nodes_where(axis, length)
    h_lo, h_hi = rng.uniform(0.0, 10.0, 2)
    bc_idx = np.concatenate([lo_face, hi_face]).astype(np.int64)
    bc_val = np.concatenate([np.full(lo_face.size, h_lo), np.full(hi_face.size, h_hi)])
    ijk = [int(rng.integers(1, n)) for _ in range(axes)]
    sink = m.node_index(*ijk)
    return Scenario(f"request_{spec.rid}_{spec.kind}{n}", m, L.LIQUID_FLOW,
                    np.full(m.n_elem, 1e-4), K, 1, bc_idx, bc_val, np.array([sink], np.int64),
                    np.array([q]), np.zeros(m.n_nodes), [Step(dt, 1.0)] * spec.steps,
                    meta=dict(seed=spec.seed, sigma_ln_k=spec.sigma_ln_k))


def bmskel() -> Scenario:
    """The reference's tests/data/bmskel deck: one line element of length 1,
    HEAT_TRANSPORT, Dirichlet 274 at POINT0 (node 0), one step of dt = 1 with
    theta = 0.5 and BiCGStab + Jacobi to 1e-12.  Solid density 2000, capacity
    900, conductivity 5 (a.msp).  The initial condition 273 applies to POINT0
    only: a.ic lists "POINT POINT0" and "POINT POINT1" under one $GEO_TYPE,
    and the reference reads the first line only (FEM/rf_ic_new.cpp:363-387),
    so node 1 starts at 0 (ogs5_amd.deck reads the deck files the same way)."""
    m = line_mesh(1, 1.0)
    return Scenario("bmskel", m, L.HEAT_TRANSPORT, np.array([2000.0 * 900.0]), np.array([5.0]), 1,
                    np.array([0], np.int64), np.array([274.0]), np.zeros(0, np.int64),
                    np.zeros(0), np.array([273.0, 0.0]), [Step(1.0, 0.5)], tol=1e-12,
                    max_iter=1000)


# ---------------------------------------------------------------------------
# GPU driver
# ---------------------------------------------------------------------------
def make_system(sc: Scenario, precision: int = L.FP64):
    pat = L.Pattern(sc.n, sc.mesh.etype, sc.mesh.conn)
    sys_ = L.System(pat, precision)
    sys_.set_elements(**sc.element_kwargs())
    return pat, sys_


def run_gpu(sc: Scenario, *, method: int = L.CG, dirichlet: int = L.DIRICHLET_SYMMETRIC,
            precision: int = L.FP64, inner_tol: Optional[float] = None, system=None,
            keep_steps=False, tol: Optional[float] = None, ir_mode: Optional[int] = None,
            ic_on_device: bool = False, fetch_x: bool = True) -> dict:
    """Runs every step of the scenario on the GPU; returns the final x, the
    per-step solve info and (keep_steps) every step's solution.

    ic_on_device: the initial condition is already in the system's stored
    vector (VEC_IC, set once by the caller) and is copied from there on the
    device, so the run makes no host-to-device transfer of node values.
    fetch_x=False leaves the final solution on the device (x is None)."""
    if system is None:
        pat, s = make_system(sc, precision)
    else:
        pat, s = system
    if ic_on_device:
        s.copy_vector(L.VEC_UOLD, L.VEC_IC)
        s.copy_vector(L.VEC_X, L.VEC_IC)
    else:
        s.set_vector(L.VEC_UOLD, sc.u0)
        s.set_vector(L.VEC_X, sc.u0)
    infos, xs = [], []
    for k, st in enumerate(sc.steps):
        _prepare_step(sc, s, k, dirichlet, ic_on_device)
        info = s.solve(method=method, tol=sc.tol if tol is None else tol,
                       max_iter=sc.max_iter, inner_tol=inner_tol, ir_mode=ir_mode)
        infos.append(info)
        s.copy_vector(L.VEC_UOLD, L.VEC_X)
        if keep_steps:
            xs.append(s.get_vector(L.VEC_X))
    x = s.get_vector(L.VEC_X) if fetch_x else None
    out = dict(x=x, info=infos, system=(pat, s))
    if keep_steps:
        out["steps"] = xs
    return out


def run_coupled_gpu(flow: Scenario, heat: Scenario, *, precision: int = L.FP64,
                    tol: Optional[float] = None, keep_vel: bool = False,
                    systems: Optional[tuple] = None) -> dict:
    """Each time step: the flow step (CG, symmetric Dirichlet), the
    Gauss-point velocity of its new pressure written into the heat system on
    the device (ogsx_gauss_velocity), then the heat step (BiCGStab, row-only
    Dirichlet: advection makes it non-symmetric).  Both systems share the
    mesh's CRS pattern.  systems: (pattern, flow system, heat system) made
    by make_coupled_systems, kept open for reuse; otherwise made and closed
    here."""
    own = systems is None
    pat, sf, sh = make_coupled_systems(flow, heat, precision) if own else systems
    for sc, s in ((flow, sf), (heat, sh)):
        s.set_vector(L.VEC_UOLD, sc.u0)
        s.set_vector(L.VEC_X, sc.u0)
    infos_f, infos_h, vel = [], [], None
    try:
        for k in range(len(flow.steps)):
            for sc, s, infos, method, dirichlet in (
                    (flow, sf, infos_f, L.CG, L.DIRICHLET_SYMMETRIC),
                    (heat, sh, infos_h, L.BICGSTAB, L.DIRICHLET_ROW)):
                _prepare_step(sc, s, k, dirichlet)
                infos.append(s.solve(method=method, tol=sc.tol if tol is None else tol,
                                     max_iter=sc.max_iter))
                s.copy_vector(L.VEC_UOLD, L.VEC_X)
                if s is sf:
                    vel = sf.gauss_velocity(sh, host=keep_vel)
        return dict(x_flow=sf.get_vector(L.VEC_X), x_heat=sh.get_vector(L.VEC_X),
                    info_flow=infos_f, info_heat=infos_h, vel=vel)
    finally:
        if own:
            sh.close()
            sf.close()


def make_coupled_systems(flow: Scenario, heat: Scenario, precision: int = L.FP64):
    pat = L.Pattern(flow.n, flow.mesh.etype, flow.mesh.conn)
    sf = L.System(pat, precision)
    sf.set_elements(**flow.element_kwargs())
    sh = L.System(pat, precision)
    sh.set_elements(**heat.element_kwargs())
    return pat, sf, sh


def _prepare_step(sc: Scenario, s, k: int, dirichlet: int, ic_on_device: bool = False):
    st = sc.steps[k]
    if st.reset_to_ic:
        if ic_on_device:
            s.copy_vector(L.VEC_UOLD, L.VEC_IC)
            s.copy_vector(L.VEC_X, L.VEC_IC)
        else:
            s.set_vector(L.VEC_UOLD, sc.u0)
            s.set_vector(L.VEC_X, sc.u0)
    s.initialize()
    s.assemble(st.dt, st.theta, st.steady)
    if sc.st_idx.size:
        s.add_rhs(sc.st_idx, sc.st_val)
    s.set_dirichlet(sc.bc_idx, sc.bc_val, dirichlet)


def step_params(sc: Scenario, k: int, dirichlet: int) -> dict:
    """Group.step parameters of step k > 0 of a scenario (not a restart from
    the initial condition): u_old <- x, then the steps of _prepare_step with
    the lists the scenario's first step uploaded."""
    st = sc.steps[k]
    return dict(dt=st.dt, theta=st.theta, steady=st.steady, dirichlet=dirichlet,
                sources=bool(sc.st_idx.size), shift=True)


def batchable(sc: Scenario, k: int) -> bool:
    """Step k can go through Group.step: not the first step (which uploads
    the source and Dirichlet lists) and not a restart from the initial
    condition."""
    return k > 0 and not sc.steps[k].reset_to_ic


def _start_steps(g, scs, systems, ids, step, dirichlet, batched, **cfg):
    """Prepares and starts step step[i] of every listed instance: through one
    Group.step for the batchable ones (batched), else on each instance's
    own stream (its copy-back u_old <- x first, for k > 0), then one group
    start per distinct (tol, max_iter)."""
    tol = cfg.pop("tol")
    by = {}
    for i in ids:
        sc = scs[i]
        by.setdefault((sc.tol if tol is None else tol, sc.max_iter), []).append(i)
    for (t, mi), grp in by.items():
        many = [i for i in grp if batched and batchable(scs[i], step[i])]
        one = [i for i in grp if i not in many]
        for i in one:
            s = systems[i][1]
            if step[i] > 0:
                s.copy_vector(L.VEC_UOLD, L.VEC_X)
            _prepare_step(scs[i], s, step[i], dirichlet)
        if one:
            g.start([systems[i][1] for i in one], tol=t, max_iter=mi, **cfg)
        if many:
            g.step([systems[i][1] for i in many],
                   [step_params(scs[i], step[i], dirichlet) for i in many], tol=t, max_iter=mi,
                   **cfg)


def run_group(scs: List[Scenario], *, method: int = L.CG,
              dirichlet: int = L.DIRICHLET_SYMMETRIC, precision: int = L.FP64,
              inner_tol: Optional[float] = None, tol: Optional[float] = None,
              ir_mode: Optional[int] = None, use_graphs: bool = True, keep_steps=False,
              stagger: int = 0, systems=None, continuous: bool = False,
              batch_iters: int = 64, batched_steps: bool = False) -> dict:
    """Runs the scenarios' steps through one ogsx group.  Lock-step (default):
    per step every instance is assembled on its own stream, then all solves
    run together.  stagger > 0 starts the instances one at a time with that
    many group iterations in between (solves join while others are
    running).  continuous: an instance assembles and starts its next step as
    soon as its solve finished, while the others keep running (each
    instance's arithmetic is its own, so results are those of lock-step).
    batched_steps: steps after the first (except restarts from the initial
    condition) are prepared for all starting instances at once
    (Group.step), with the same A, b and results."""
    if systems is None:
        systems = [make_system(sc, precision) for sc in scs]
    g = L.Group(precision, use_graphs)
    for sc, (_, s) in zip(scs, systems):
        g.add(s)
        s.set_vector(L.VEC_UOLD, sc.u0)
        s.set_vector(L.VEC_X, sc.u0)
    infos = [[] for _ in scs]
    xs = [[] for _ in scs]
    cfg = dict(method=method, tol=tol, inner_tol=inner_tol, ir_mode=ir_mode)
    if continuous:
        return _run_group_continuous(scs, systems, g, infos, xs, dirichlet, keep_steps,
                                     batch_iters, batched_steps, cfg)
    for k in range(max(len(sc.steps) for sc in scs)):
        live = [i for i, sc in enumerate(scs) if k < len(sc.steps)]
        step = [k] * len(scs)
        if stagger:
            for i in live:
                _start_steps(g, scs, systems, [i], step, dirichlet, batched_steps, **cfg)
                g.advance(stagger)
        else:
            _start_steps(g, scs, systems, live, step, dirichlet, batched_steps, **cfg)
        while g.advance() > 0:
            pass
        for i in live:
            s = systems[i][1]
            active, info = L.solve_status(s)
            assert not active
            infos[i].append(info)
            if keep_steps:
                xs[i].append(s.get_vector(L.VEC_X))
    out = dict(x=[s.get_vector(L.VEC_X) for _, s in systems], info=infos, systems=systems,
               group=g)
    if keep_steps:
        out["steps"] = xs
    return out


def _run_group_continuous(scs, systems, g, infos, xs, dirichlet, keep_steps, batch_iters,
                          batched_steps, cfg):
    step = [0] * len(scs)

    def start(ids):
        _start_steps(g, scs, systems, ids, step, dirichlet, batched_steps, **cfg)

    live = [i for i, sc in enumerate(scs) if sc.steps]
    start(live)
    while live:
        g.advance(batch_iters)
        nxt, still = [], []
        for i in live:
            s = systems[i][1]
            active, info = L.solve_status(s)
            if active:
                still.append(i)
                continue
            infos[i].append(info)
            if keep_steps:
                xs[i].append(s.get_vector(L.VEC_X))
            step[i] += 1
            if step[i] < len(scs[i].steps):
                nxt.append(i)
                still.append(i)
        if nxt:
            start(nxt)
        live = still
    out = dict(x=[s.get_vector(L.VEC_X) for _, s in systems], info=infos, systems=systems,
               group=g)
    if keep_steps:
        out["steps"] = xs
    return out
