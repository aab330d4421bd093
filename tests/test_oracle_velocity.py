"""The oracle's Gauss-point Darcy velocity (Cal_Velocity) pinned by fields
the elements represent exactly.

* a linear pressure on distorted elements (isoparametric elements reproduce
  it): v = -(K/mu)(grad p + rho_g e_dim) at every Gauss point, with a
  non-symmetric K so that a transposed tensor, a dropped or mis-signed
  gravity term, or a lost tuf/mu factor fails;
* a bilinear (trilinear) pressure on axis-aligned boxes, whose gradient
  (y, x) / (yz, xz, xy) varies from point to point: this pins the Gauss
  point order against the reference's own shape-function values at its
  Gauss points (tests/golden/ref_mathlib_dump.txt);
* the coupled flow -> heat driver on a homogeneous block, whose steady
  pressure is linear: the velocity is the closed-form Darcy flux.
"""
import os

import numpy as np
import pytest

from ogs5_amd import mesh as M
from ogs5_amd import workloads as W
from oracle import driver

LIQ = 1
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ref_mathlib_dump.txt")


def _coords(m):
    cols = [m.x, m.y, m.z][: m.dim]
    return np.stack(cols, 1)


def _distorted(kind, rng):
    if kind == "line":
        m = M.line_mesh(5, 3.0)
        m.x = m.x + np.r_[0.0, 0.2 * rng.uniform(-1, 1, m.n_nodes - 2), 0.0]
    elif kind == "quad":
        m = M.quad_mesh(4, 3, 2.0, 1.5)
        m.x = m.x + 0.08 * rng.uniform(-1, 1, m.n_nodes)
        m.y = m.y + 0.08 * rng.uniform(-1, 1, m.n_nodes)
    elif kind == "hex":
        m = M.hex_mesh(3, 2, 2)
        for a in ("x", "y", "z"):
            setattr(m, a, getattr(m, a) + 0.12 * rng.uniform(-1, 1, m.n_nodes))
    else:
        m = M.tet_mesh(2, 2, 2)
        A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        X = np.stack([m.x, m.y, m.z], 1) @ A.T
        m.x, m.y, m.z = (X[:, i].copy() for i in range(3))
    return m


@pytest.mark.parametrize("kind", ["line", "quad", "hex", "tet"])
@pytest.mark.parametrize("gravity", [False, True])
def test_linear_pressure_gives_darcy_flux(ffi, kind, gravity):
    rng = np.random.default_rng(11)
    m = _distorted(kind, rng)
    D = m.dim
    K = (np.eye(D) + 0.4 * rng.uniform(-1, 1, (D, D))) * 3e-12  # not symmetric
    mu, tuf, rho_g = 1.3e-3, 86400.0, (9.81e3 if gravity else 0.0)
    a = rng.uniform(0.5, 2.0, D) * 1e4 * rng.choice([-1, 1], D)
    u = _coords(m) @ a + 7.0
    p = ffi.OracleProblem(m, process=LIQ, mass_coef=np.ones(m.n_elem),
                          cond=np.tile(K.ravel(), m.n_elem), cond_stride=D * D, tuf=tuf, mu=mu,
                          rho_g=rho_g)
    vel = p.gauss_velocity(u)
    g = a.copy()
    g[D - 1] += rho_g
    expect = -(K / mu) @ g
    got = vel[:, :D, :]
    tol = 1e-11 * np.abs(expect).max()
    assert np.abs(got - expect[None, :, None]).max() < tol
    assert not vel[:, D:, :].any()
    if D > 1:  # the transposed tensor would be a different flux
        assert np.abs(-(K.T / mu) @ g - expect).max() > 1e3 * tol


def _golden_shapes(etype):
    rows = {}
    for ln in open(GOLDEN).read().splitlines():
        t = ln.split()
        if t[0] == "shape" and int(t[1]) == etype:
            rows.setdefault(int(t[2]), {})[int(t[3])] = [float.fromhex(v) for v in t[4:]]
    return rows


@pytest.mark.parametrize("kind", ["quad", "hex"])
def test_gauss_point_order_matches_reference(ffi, kind):
    if kind == "quad":
        m, etype, npe = M.quad_mesh(3, 2, 3.0, 1.0), 2, 4
        field = lambda X: X[:, 0] * X[:, 1]
        grad = lambda P: np.stack([P[:, 1], P[:, 0]], 1)
    else:
        m, etype, npe = M.hex_mesh(2, 2, 1, 2.0, 1.0, 0.5), 3, 8
        field = lambda X: X[:, 0] * X[:, 1] * X[:, 2]
        grad = lambda P: np.stack([P[:, 1] * P[:, 2], P[:, 0] * P[:, 2], P[:, 0] * P[:, 1]], 1)
    D = m.dim
    X = _coords(m)
    K = np.diag(np.arange(1.0, D + 1.0)) * 1e-11
    mu = 1e-3
    shapes = _golden_shapes(etype)
    checked = 0
    for ng, pts in shapes.items():
        if ng < 2:
            continue
        p = ffi.OracleProblem(m, process=LIQ, mass_coef=np.ones(m.n_elem),
                              cond=np.tile(K.ravel(), m.n_elem), cond_stride=D * D, mu=mu,
                              n_gauss=ng)
        vel = p.gauss_velocity(field(X))
        for gp, vals in pts.items():
            N = np.array(vals[:npe])
            P = np.einsum("a,ead->ed", N, X[m.conn])  # Gauss point of every element
            expect = -(grad(P) @ (K / mu).T)
            np.testing.assert_allclose(vel[:, :D, gp], expect, rtol=1e-12,
                                       atol=1e-13 * np.abs(expect).max())
            checked += 1
    assert checked >= 4 + 9


def test_coupled_steady_homogeneous_flux(ffi):
    flow, heat = W.coupled_flow_heat(4, sigma_ln_k=0.0, n_steps=2)
    flow.steps = [W.Step(s.dt, s.theta, steady=True) for s in flow.steps]
    flow.tol = heat.tol = 1e-14
    out = driver.run_oracle_coupled(flow, heat, ffi)
    k = flow.cond[0]
    assert np.all(flow.cond == k)
    lx = float(flow.mesh.x.max())
    vx = k / flow.mu * 2.0e5 / lx
    v = out["vel"]
    assert np.abs(v[:, 0, :] - vx).max() < 1e-9 * vx
    assert np.abs(v[:, 1:, :]).max() < 1e-9 * vx
    # heat entered through the inlet face (Galerkin at this cell Peclet
    # number over- and undershoots, so no bounds are asserted)
    T = out["x_heat"]
    assert np.abs(T[heat.bc_idx] - 60.0).max() < 1e-8 and T.mean() > 20.0
    assert all(i["error"] <= 1e-14 for i in out["info_flow"] + out["info_heat"])
