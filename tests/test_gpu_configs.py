"""The batched-ensemble and served-stream configurations at their full
sizes, and failure isolation inside a group and the server.

* 256 two-dimensional flow instances (quad side 64-512 log-uniform,
  sigma_lnK 0.5-2.5, 10 implicit steps each) in one FP32-inside-FP64 group:
  every solve reaches a true FP64 relative residual <= 1e-10; sampled
  instances are bit-identical to the same instance solved alone; the
  smallest ones solve the oracle's own assembled systems (residual against
  the oracle's matrix) and, re-solved tight, match the oracle's tight chain
  within 1e-10 relative L2.
* 100 requests of the served stream (Poisson 200/s, quads 32^2-512^2 and
  hexes 16^3-96^3, 1-20 steps): all converge, sampled results equal the
  request solved alone, the small ones are checked against the oracle as
  above; p50/p99 latency printed.
* A member that hits its iteration limit, or carries a NaN in its
  right-hand side, reports converged = 0 while the members around it stay
  bit-identical to alone (the reference prints a warning and returns,
  FEM/equation_class.cpp:1092-1093; a NaN solution rejects the step,
  FEM/rf_pcs.cpp:5150-5153).
"""
import numpy as np
import pytest

import ogs5_amd as ox
from ogs5_amd import _lib as L
from ogs5_amd import workloads as W
from ogs5_amd.serve import Server
from oracle import driver

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _oracle_residuals(sc, xs, ffi):
    """||b - A x_k|| / ||b|| of each step's device solution x_k on the
    oracle's own system of that step (assembled from x_{k-1}, sources,
    symmetric Dirichlet elimination)."""
    ptr, col, diag = ffi.pattern(sc.n, sc.mesh.conn)
    eq = ffi.OracleEqs(ptr, col, diag)
    prob = ffi.OracleProblem(sc.mesh, process=sc.process, mass_coef=sc.mass_coef, cond=sc.cond,
                             cond_stride=sc.cond_stride, n_gauss=sc.n_gauss, tuf=sc.tuf,
                             mu=sc.mu, geo_area=sc.geo_area, adv_coef=sc.adv_coef, vel=sc.vel,
                             rho_g=sc.rho_g, u_old=sc.u0)
    out, u_old = [], sc.u0.copy()
    for st, x in zip(sc.steps, xs):
        if st.reset_to_ic:
            u_old = sc.u0.copy()
        prob.set(dt=st.dt, theta=st.theta, steady=st.steady, u_old=u_old)
        eq.initialize()
        eq.assemble(prob)
        if sc.st_idx.size:
            eq.add_rhs(sc.st_idx, sc.st_val)
        eq.set_known_x(sc.bc_idx, sc.bc_val, symmetric=True)
        b = eq.b.copy()
        out.append(np.linalg.norm(b - eq.spmv(x)) / np.linalg.norm(b))
        u_old = x
    return out


def _tight_vs_oracle(scs, ffi):
    """The instances re-solved in one FP32-IR group to 1e-13 against the
    oracle's FP64 CG chain at 1e-14: max relative L2 error over all steps."""
    mix = W.run_group(scs, precision=ox.FP32_IR, tol=1e-13, keep_steps=True)
    worst = 0.0
    for k, sc in enumerate(scs):
        ref = driver.run_oracle(sc, ffi, method=5, symmetric=True, tol=1e-14, keep_steps=True)
        for x, xr in zip(mix["steps"][k], ref["steps"]):
            worst = max(worst, np.linalg.norm(x - xr) / np.linalg.norm(xr))
    return worst


def _same_as_alone(sc, x, infos, prec):
    alone = W.run_gpu(sc, precision=prec, tol=TOL, keep_steps=True)
    assert np.array_equal(alone["x"], x), sc.name
    assert [i["iterations"] for i in alone["info"]] == [i["iterations"] for i in infos]
    assert [i["error"] for i in alone["info"]] == [i["error"] for i in infos]
    return alone["steps"]


def test_ensemble_256_instances_fp32_ir_group(ffi):
    scs = W.ensemble(256)
    dofs = np.array([sc.n for sc in scs])
    print(f"[ensemble] 256 instances, n {dofs.min()}-{dofs.max()}, sum {dofs.sum()}",
          flush=True)
    mix = W.run_group(scs, precision=ox.FP32_IR, tol=TOL)
    errs = [i["error"] for infos in mix["info"] for i in infos]
    assert len(errs) == 256 * 10
    assert all(i["converged"] == 1 for infos in mix["info"] for i in infos)
    assert max(errs) <= TOL  # the IR stopping test is the true FP64 residual
    print(f"[ensemble] true FP64 residual max {max(errs):.3e}", flush=True)
    mix["group"].close()
    x_group, info_group = mix["x"], mix["info"]
    for _, s in mix["systems"]:
        s.close()
    rng = np.random.default_rng(2024)
    smallest = list(np.argsort(dofs, kind="stable")[:3])
    sample = sorted(set(smallest) | set(rng.choice(256, 5, replace=False).tolist()))
    for k in sample:
        steps = _same_as_alone(scs[k], x_group[k], info_group[k], ox.FP32_IR)
        if k in smallest:
            res = _oracle_residuals(scs[k], steps, ffi)
            print(f"[ensemble] instance {k} n={scs[k].n}: residual on the oracle's systems "
                  f"max {max(res):.3e}", flush=True)
            assert max(res) <= 1.5 * TOL
    worst = _tight_vs_oracle([scs[k] for k in smallest], ffi)
    print(f"[ensemble] smallest three re-solved to 1e-13: max rel L2 vs oracle {worst:.3e}",
          flush=True)
    assert worst < 1e-10


def test_served_stream_at_200_per_s(ffi):
    specs = W.request_stream(100, rate=200.0, seed=11)
    srv = Server(ox.FP32_IR)
    out = srv.serve(specs, keep_x=True)
    assert out["requests"] == 100 and not out["not_converged"]
    lat = out["latency_ms"]
    print(f"[serve] 100 requests at 200/s: {out['requests_per_s']:.1f} req/s, "
          f"{out['mdof_solves_per_s']:.1f} MDOF-solves/s, p50 {lat['p50']:.0f} ms, "
          f"p99 {lat['p99']:.0f} ms", flush=True)
    jobs = sorted(out["jobs"], key=lambda j: j.sc.n)
    small = jobs[:3]
    rng = np.random.default_rng(7)
    others = [jobs[i] for i in rng.choice(np.arange(3, len(jobs)), 3, replace=False)]
    for job in small + others:
        steps = _same_as_alone(job.sc, job.x, job.infos, ox.FP32_IR)
        if job in small:
            res = _oracle_residuals(job.sc, steps, ffi)
            print(f"[serve] request {job.spec.rid} ({job.spec.kind}{job.spec.side}, "
                  f"{job.spec.steps} steps): residual on the oracle's systems max "
                  f"{max(res):.3e}", flush=True)
            assert max(res) <= 1.5 * TOL
    worst = _tight_vs_oracle([j.sc for j in small], ffi)
    print(f"[serve] three smallest re-solved to 1e-13: max rel L2 vs oracle {worst:.3e}",
          flush=True)
    assert worst < 1e-10


def _failing_mix():
    scs = [W.quad_flow(n, n, seed=50 + n, sigma_ln_k=1.5, n_steps=2) for n in (12, 30, 45, 20)]
    scs[1].max_iter = 5           # hits its iteration limit
    scs[2].st_val = np.array([np.nan])  # NaN source: NaN right-hand side
    scs[2].max_iter = 40
    return scs


@pytest.mark.parametrize("prec", [ox.FP64, ox.FP32_IR])
def test_group_failure_isolation(prec):
    scs = _failing_mix()
    mix = W.run_group(scs, precision=prec, tol=TOL, keep_steps=True)
    for k in (0, 3):
        alone = W.run_gpu(scs[k], precision=prec, tol=TOL, keep_steps=True)
        for xa, xb in zip(alone["steps"], mix["steps"][k]):
            assert np.array_equal(xa, xb)
        assert [i["iterations"] for i in alone["info"]] == \
            [i["iterations"] for i in mix["info"][k]]
        assert all(i["converged"] == 1 for i in mix["info"][k])
    assert all(i["converged"] == 0 for i in mix["info"][1])
    assert all(i["iterations"] <= 5 for i in mix["info"][1])
    assert all(i["converged"] == 0 for i in mix["info"][2])
    assert np.isnan(mix["steps"][2][-1]).any()
    print(f"[isolation] prec={prec} limited {[i['iterations'] for i in mix['info'][1]]} "
          f"nan {[i['iterations'] for i in mix['info'][2]]}", flush=True)


def test_server_lists_failed_requests(monkeypatch):
    specs = W.request_stream(12, rate=400.0, seed=5, quad_sides=(8, 40), hex_sides=(4, 10),
                             steps=(1, 3))
    bad_limit, bad_nan = specs[3].rid, specs[7].rid
    make = W.request_scenario

    def scenario(spec, mesh=None):
        sc = make(spec, mesh)
        if spec.rid == bad_limit:
            sc.max_iter = 3
        if spec.rid == bad_nan:
            sc.st_val = np.array([np.nan])
            sc.max_iter = 30
        return sc

    monkeypatch.setattr(W, "request_scenario", scenario)
    srv = Server(ox.FP32_IR, max_active=6, batch_iters=8, workers=4)
    out = srv.serve(specs, keep_x=True)
    assert out["requests"] == 12
    assert sorted(out["not_converged"]) == sorted([bad_limit, bad_nan])
    for job in out["jobs"]:
        if job.spec.rid in (bad_limit, bad_nan):
            assert all(i["converged"] == 0 for i in job.infos)
            continue
        alone = W.run_gpu(job.sc, precision=ox.FP32_IR, tol=TOL)
        assert np.array_equal(alone["x"], job.x), job.spec
