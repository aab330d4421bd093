"""The C++ host layer (include/ogsx.hpp) driving a transient LIQUID_FLOW run
in the reference's call order, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from ogs5_amd import workloads as W
from oracle import driver

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "ogs5_amd", "process_step")


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
