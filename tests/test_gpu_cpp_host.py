"""The C++ host layer (include/ogsx.hpp) driving a transient LIQUID_FLOW run,
and a coupled LIQUID_FLOW -> Gauss-point velocity -> HEAT_TRANSPORT run, in
the reference's call order, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from ogs5_amd import workloads as W
from oracle import driver

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ogs5_amd", "process_step")
EXE_COUPLED = os.path.join(ROOT, "ogs5_amd", "coupled_step")


def _write_input(path, sc, method, symmetric, tol):
    m = sc.mesh
    st = sc.steps[0]
    with open(path, "wb") as f:
        np.array([m.n_nodes, m.n_elem, m.etype, sc.process, m.dim, len(sc.steps), method,
                  int(symmetric)], np.int64).tofile(f)
        np.array([st.dt, st.theta, tol], np.float64).tofile(f)
        for a in (m.x, m.y, m.z):
            np.asarray(a, np.float64).tofile(f)
        np.asarray(m.conn, np.int64).tofile(f)
        np.asarray(sc.mass_coef, np.float64).tofile(f)
        np.asarray(sc.cond, np.float64).tofile(f)
        np.array([sc.bc_idx.size], np.int64).tofile(f)
        sc.bc_idx.astype(np.int64).tofile(f)
        sc.bc_val.astype(np.float64).tofile(f)
        np.array([sc.st_idx.size], np.int64).tofile(f)
        sc.st_idx.astype(np.int64).tofile(f)
        sc.st_val.astype(np.float64).tofile(f)
        sc.u0.astype(np.float64).tofile(f)


@pytest.mark.parametrize("method,symmetric", [(2, False), (5, True)])
def test_cpp_process_step_vs_oracle(tmp_path, method, symmetric):
    from oracle import ffi
    sc = W.quad_flow(30, 24, seed=9, n_steps=3)
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_input(inp, sc, method, symmetric, 1e-13)
    print(f"[cpp] method={method} n={sc.n}", flush=True)
    subprocess.run([EXE, str(inp), str(out)], check=True, timeout=120)
    raw = np.fromfile(out, np.float64, count=len(sc.steps) * sc.n).reshape(len(sc.steps), sc.n)
    ref = driver.run_oracle(sc, ffi, method=5, symmetric=True, tol=1e-14, keep_steps=True)
    for x, xr in zip(raw, ref["steps"]):
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-10


def test_cpp_coupled_step_vs_oracle(tmp_path):
    from oracle import ffi
    flow, heat = W.coupled_flow_heat(4, n_steps=3)
    m = flow.mesh
    st = heat.steps[0]
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([m.n_nodes, m.n_elem, m.etype, m.dim, len(flow.steps)], np.int64).tofile(f)
        np.array([st.dt, st.theta, 1e-13, flow.mu, flow.rho_g, heat.adv_coef],
                 np.float64).tofile(f)
        for a in (m.x, m.y, m.z):
            np.asarray(a, np.float64).tofile(f)
        np.asarray(m.conn, np.int64).tofile(f)
        for sc in (flow, heat):
            np.asarray(sc.mass_coef, np.float64).tofile(f)
            np.asarray(sc.cond, np.float64).tofile(f)
            np.array([sc.bc_idx.size], np.int64).tofile(f)
            sc.bc_idx.astype(np.int64).tofile(f)
            sc.bc_val.astype(np.float64).tofile(f)
            sc.u0.astype(np.float64).tofile(f)
    print(f"[cpp] coupled n={m.n_nodes} E={m.n_elem}", flush=True)
    subprocess.run([EXE_COUPLED, str(inp), str(out)], check=True, timeout=120)
    n, E = m.n_nodes, m.n_elem
    raw = np.fromfile(out, np.float64, count=2 * n + E * 3 * 5)
    p, T, vel = raw[:n], raw[n:2 * n], raw[2 * n:].reshape(E, 3, 5)
    ref = driver.run_oracle_coupled(flow, heat, ffi, tol=1e-14)
    for got, want in ((p, ref["x_flow"]), (T, ref["x_heat"])):
        assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-10
    assert np.abs(vel - ref["vel"]).max() <= 1e-8 * np.abs(ref["vel"]).max()
