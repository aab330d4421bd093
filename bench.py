"""Benchmark of the implicit FEM step on the 3-D liquid-flow workload.

Workload: 128^3 trilinear hexes (2,146,689 DOF, 57,066,625 nonzeros),
lognormal K (sigma_lnK = 2, 8-cell correlation), head 1 m across the x faces.
One bench step is the whole job of that configuration: a steady solve plus
5 implicit-Euler transient steps, each = Initialize + element assembly +
Dirichlet + Jacobi-PCG to ||r||/||b|| <= 1e-10 (true FP64 residual).  It is
run twice per step: plain FP64 and FP32 matrix/SpMV/preconditioner inside
FP64 iterative refinement.  Inputs (mesh, K field, pattern, initial
condition) are uploaded to HBM before the timed region; the jobs copy the
initial condition on the device and leave the solution there.  The same jobs
with u0 and x crossing PCIe are timed after them (`pcie_inclusive`).

Prints one JSON line.  value = MDOF-solves/s of the faster valid solver
(both reach 1e-10 FP64 residual); the plain FP64 figures ride along.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def run_job(sc, system, precision, resident=True):
    """One job.  resident: the initial condition is copied from the system's
    stored vector on the device and the solution stays there (no PCIe
    transfer of node values); otherwise u0 is uploaded from the host at the
    start and at the transient restart and x is downloaded at the end."""
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    res = W.run_gpu(sc, method=ox.CG, dirichlet=ox.DIRICHLET_SYMMETRIC, precision=precision,
                    system=system, ic_on_device=resident, fetch_x=not resident)
    return res["info"]


def gpu_leg(sc, precision, steps, warmup):
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    system = W.make_system(sc, precision)
    pat, s = system
    s.set_vector(ox.VEC_IC, sc.u0)   # input resident in HBM before any timing
    for _ in range(warmup):
        info = run_job(sc, system, precision)
        print(f"[bench] warmup precision={precision} iters={[i['iterations'] for i in info]}",
              file=sys.stderr, flush=True)
    s.synchronize()
    t0 = time.perf_counter()
    infos = []
    for _ in range(steps):
        infos.append(run_job(sc, system, precision))
    s.synchronize()
    dt = time.perf_counter() - t0
    print(f"[bench] timed precision={precision} steps={steps} {dt:.3f} s", file=sys.stderr,
          flush=True)
    # the same jobs with the node values crossing PCIe (u0 in, x out), as a
    # caller holding them in host memory would run them; reported beside
    # the resident figure, never as the value
    t0 = time.perf_counter()
    for _ in range(steps):
        run_job(sc, system, precision, resident=False)
    s.synchronize()
    dt_pcie = time.perf_counter() - t0
    # one more job, untimed, with HIP events around every SpMV launch (the
    # event records add a few microseconds of idle time per launch, so they
    # stay out of the timed region)
    s.profile(True)
    steps_x = W.run_gpu(sc, method=ox.CG, dirichlet=ox.DIRICHLET_SYMMETRIC, precision=precision,
                        system=system, ic_on_device=True, fetch_x=False, keep_steps=True)["steps"]
    kind = ox._lib.PROF_SPMV64 if precision == ox.FP64 else ox._lib.PROF_SPMV32
    ms, launches, bpl = s.profile_read(kind)
    ms64, l64, b64 = s.profile_read(ox._lib.PROF_SPMV64)
    s.profile(False)
    # the dominant kernel alone: the solve's SpMV kernel (k_cg_spmv, tuned
    # format) on the last step's matrix, 50 launches timed with HIP events on
    # the system's stream.  The in-solve average above also counts launches
    # that found the recurrence already stopped (early exit within a batch).
    k_ms, k_bytes = s.bench_spmv(precision == ox.FP32_IR, 50)
    solves = len(sc.steps) * steps
    last = infos[-1]
    spmv_in_ms = ms / max(launches, 1)
    assert all(i["converged"] == 1 for job in infos for i in job), infos[-1]
    assert all(i["error"] <= sc.tol for job in infos for i in job)
    return dict(
        seconds=dt, ms_per_step=1e3 * dt / steps,
        mdof_solves_per_s=solves * sc.n / dt / 1e6,
        pcie_inclusive=dict(ms_per_step=1e3 * dt_pcie / steps,
                            mdof_solves_per_s=solves * sc.n / dt_pcie / 1e6),
        iterations=[i["iterations"] for i in last], outer=[i["outer"] for i in last],
        errors=[i["error"] for i in last],
        spmv=dict(avg_ms=k_ms, bytes_per_launch=k_bytes, gbs=k_bytes / k_ms / 1e6, launches=50,
                  kernel_alone=True),
        spmv_in_solve=dict(avg_ms=spmv_in_ms, launches=launches, bytes_per_launch=bpl,
                           gbs=bpl / spmv_in_ms / 1e6 if launches else None,
                           frac=bpl / spmv_in_ms / 1e6 / HBM_PEAK_GBS if launches else None,
                           iterations=sum(i["iterations"] for i in infos[-1]),
                           untimed_job=True),
        spmv64=dict(avg_ms=ms64 / max(l64, 1), launches=l64, bytes_per_launch=b64,
                    gbs=b64 / (ms64 / max(l64, 1)) / 1e6 if l64 else None),
        spmv_variant=s.spmv_variant(),
        spmv_format=s.spmv_format(precision == ox.FP32_IR),
        spmv_tuning=s.spmv_tuning(),
        steps_x=steps_x,
        system=system,
    )


def error_vs_fp64_ref(sc, system, legs, tol_ref=1e-13, method=None, dirichlet=None,
                      with_ref=False):
    """Relative L2 error of every step's solution of each leg against the
    same job solved in FP64 to a tight residual (the device FP64 solver of
    the job -- Jacobi-CG by default -- whose solutions match the oracle within
    1e-10 at every size the oracle can run).  Errors include what earlier
    steps pass on through u_old.  with_ref: returns (errors, the reference's
    step solutions)."""
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    method = ox.CG if method is None else method
    dirichlet = ox.DIRICHLET_SYMMETRIC if dirichlet is None else dirichlet
    ref = W.run_gpu(sc, method=method, dirichlet=dirichlet, precision=ox.FP64,
                    system=system, ic_on_device=True, fetch_x=False, keep_steps=True,
                    tol=tol_ref)
    assert all(i["converged"] == 1 for i in ref["info"]), ref["info"]
    name = "Jacobi-CG" if method == ox.CG else "Jacobi-BiCGStab"
    out = dict(reference=f"FP64 {name} job to {tol_ref:g} relative residual on the device",
               reference_residuals=[i["error"] for i in ref["info"]])
    for name, xs in legs.items():
        out[name] = [float(np.linalg.norm(x - r) / np.linalg.norm(r))
                     for x, r in zip(xs, ref["steps"])]
    return (out, ref["steps"]) if with_ref else out


def cpu_baseline(sc, iters_per_solve):
    """Serial oracle (the CPU restatement of the reference path) on a bounded
    sample of the job itself: one full assembly of the 128^3 steady
    operator, the reference's row-only Dirichlet elimination (Diagonize)
    and 20 Jacobi-CG iterations on that matrix, timed; the job's CPU time is
    taken as one assembly per solve plus the iteration time times the
    iteration counts the GPU's FP64 solve needed (same algorithm, same
    tolerance).  The CRS pattern is built before timing, as on the GPU."""
    from oracle import ffi
    ptr, col, diag = ffi.pattern(sc.n, sc.mesh.conn)
    eq = ffi.OracleEqs(ptr, col, diag)
    st = sc.steps[0]
    prob = ffi.OracleProblem(sc.mesh, process=sc.process, mass_coef=sc.mass_coef, cond=sc.cond,
                             dt=st.dt, theta=st.theta, steady=st.steady, u_old=sc.u0)
    t0 = time.perf_counter()
    eq.initialize()
    eq.assemble(prob)
    eq.set_known_x(sc.bc_idx, sc.bc_val, symmetric=False)
    t_asm = time.perf_counter() - t0
    its = 20
    eq.config(5, 1, 1e-30, its)
    t0 = time.perf_counter()
    out = eq.solve()
    t_it = (time.perf_counter() - t0) / max(out["iterations"], 1)
    n_solves = len(iters_per_solve)
    total = n_solves * t_asm + sum(iters_per_solve) * t_it
    return dict(value=n_solves * sc.n / total / 1e6, unit="MDOF-solves/s", cores=1,
                nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)), kind="port",
                sample=(f"oracle (serial C, 1 thread) on the {sc.mesh.shape[0]}^3 job itself: "
                        f"assembly + row-only Dirichlet {t_asm:.2f} s for {sc.mesh.n_elem} hexes, "
                        f"{out['iterations']} Jacobi-CG iterations at {1e3 * t_it:.1f} ms on its "
                        f"{int(ptr[-1])} nonzeros; job time = {n_solves} assemblies + "
                        f"{sum(iters_per_solve)} iterations (the GPU FP64 solve's counts) = "
                        f"{total:.1f} s"),
                assembly_s=t_asm, iteration_s=t_it, job_s=total)


def spmv_mode(size):
    """Short run for hardware-counter passes: a 1 GiB copy (calibration of
    the byte counters) and 20 launches each of the solve's SpMV kernel
    (k_cg_spmv, tuned format) on the FP64 and FP32 matrix of the assembled
    steady 128^3 system."""
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    sc = W.hex_flow(size, n_transient=0)
    pat, s = W.make_system(sc, ox.FP32_IR)
    st = sc.steps[0]
    s.set_vector(ox.VEC_UOLD, sc.u0)
    s.set_vector(ox.VEC_X, sc.u0)
    s.initialize()
    s.assemble(st.dt, st.theta, True)
    s.set_dirichlet(sc.bc_idx, sc.bc_val, ox.DIRICHLET_SYMMETRIC)
    copy = ox._lib.stream_copy_gbs(1 << 30, 5)
    ms64, b64 = s.bench_spmv(False, 20)
    ms32, b32 = s.bench_spmv(True, 20)
    print(json.dumps({"copy_1GiB_gbs": copy, "spmv64": {"avg_ms": ms64, "bytes": b64,
                                                          "gbs": b64 / ms64 / 1e6},
                      "spmv32": {"avg_ms": ms32, "bytes": b32, "gbs": b32 / ms32 / 1e6},
                      "format64": s.spmv_format(False), "format32": s.spmv_format(True),
                      "tuning": s.spmv_tuning()}))


def assembly_mode():
    """Assembly alone, for kernel traces and hardware-counter passes: a 1 GiB
    copy (byte-counter calibration), then per system one warm-up and 10
    timed assemblies (Initialize + element pass + row pass, HIP events on
    the system's stream) of the 128^3 hex flow system (transient step), the
    96^3 x 6 tet heat system (advection, Crank-Nicolson) and the 512^2 quad
    flow system.  Algorithmic bytes per assembly: per element its
    connectivity, coefficients and (heat) Gauss-point velocities read once,
    per node its coordinates and u_old read once, A and b written once."""
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    out = {"copy_1GiB_gbs": ox._lib.stream_copy_gbs(1 << 30, 5)}
    for name, sc in (("hex128_flow", W.hex_flow(128, n_transient=1)),
                     ("tet96_heat", W.tet_heat(96, n_steps=1)),
                     ("quad512_flow", W.quad_flow(512, 512, n_steps=1))):
        pat, s = W.make_system(sc)
        st = sc.steps[-1]
        s.set_vector(ox.VEC_UOLD, sc.u0)
        s.initialize()
        s.assemble(st.dt, st.theta, False)
        s.synchronize()
        s.profile(True)
        for _ in range(10):
            s.initialize()
            s.assemble(st.dt, st.theta, False)
        ms, launches, _ = s.profile_read(ox._lib.PROF_ASSEMBLY)
        s.profile(False)
        m = sc.mesh
        npe = m.conn.shape[1]
        per_elem = 4 * npe + 8 + 8 * sc.cond_stride
        if sc.vel is not None:
            per_elem += 8 * sc.vel[0].size
        alg = m.n_elem * per_elem + sc.n * (8 * m.dim + 8) + 8 * s.nnz + 8 * sc.n
        per = ms / max(launches, 1)
        out[name] = dict(elements=int(m.n_elem), nodes=int(sc.n), nnz=int(s.nnz),
                         ms_per_assembly=per, algorithmic_bytes=int(alg),
                         gbs=alg / per / 1e6)
        _log(f"assembly {name}: {1e3 * per:.0f} us per assembly, {alg / per / 1e6:.0f} GB/s "
             f"algorithmic")
        s.close()
    print(json.dumps(out))


def _log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def ensemble_runs(count, steps, group_fp64=True):
    """Batched ensemble: `count` independent 2-D flow instances (quad side
    log-uniform 64-512, sigma_lnK 0.5-2.5), `steps` implicit steps each.
    Timed runs of the whole job: one group in FP32-IR (and, with
    group_fp64, in FP64) -- one launch per Krylov phase over all running
    instances, iteration batches replayed as HIP graphs -- and the plain
    method (ogsx_solve on one instance at a time, FP64).  Untimed warm-up:
    every instance's first step (SpMV format tuning, graph capture).
    Returns (runs, scenarios)."""
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    ox.device_count()
    scs = W.ensemble(count, n_steps=steps)
    total = sum(sc.n for sc in scs) * steps
    warm = [W.Scenario(**{**sc.__dict__, "steps": sc.steps[:1]}) for sc in scs]
    res = {}
    legs = [("group_fp32_ir", ox.FP32_IR, True)]
    legs += [("group_fp64", ox.FP64, True)] if group_fp64 else [("plain_fp64", ox.FP64, False)]
    for name, prec, grouped in legs:
        systems = [W.make_system(sc, prec) for sc in scs]
        if grouped:
            W.run_group(warm, precision=prec, systems=systems)
            for _, s in systems:
                s.synchronize()
            t0 = time.perf_counter()
            out = W.run_group(scs, precision=prec, systems=systems)
            for _, s in systems:
                s.synchronize()
            dt = time.perf_counter() - t0
            its = [i["iterations"] for inf in out["info"] for i in inf]
            assert all(i["converged"] == 1 and i["error"] <= 1e-10
                       for inf in out["info"] for i in inf)
            res[name] = dict(seconds=dt, mdof_solves_per_s=total / dt / 1e6,
                             iterations_max=max(its), iterations_mean=float(np.mean(its)))
            _log(f"ensemble {name}: {dt:.2f} s, {total / dt / 1e6:.1f} MDOF-solves/s")
            out["group"].close()
        if prec == ox.FP64:
            for sc, sysw in zip(warm, systems):
                W.run_gpu(sc, precision=ox.FP64, system=sysw)
            for _, s in systems:
                s.synchronize()
            t0 = time.perf_counter()
            its = []
            for sc, sysw in zip(scs, systems):
                r = W.run_gpu(sc, precision=ox.FP64, system=sysw)
                its += [i["iterations"] for i in r["info"]]
                assert all(i["converged"] == 1 and i["error"] <= 1e-10 for i in r["info"])
            dt = time.perf_counter() - t0
            res["plain_fp64"] = dict(seconds=dt, mdof_solves_per_s=total / dt / 1e6,
                                     iterations_max=max(its), iterations_mean=float(np.mean(its)))
            _log(f"ensemble plain_fp64: {dt:.2f} s, {total / dt / 1e6:.1f} MDOF-solves/s")
        for _, s in systems:
            s.close()
        del systems
    return res, scs


def ensemble_mode(count, steps):
    res, scs = ensemble_runs(count, steps)
    best = res["group_fp32_ir"]
    print(json.dumps({
        "metric": "MDOF-solves/s to 1e-10 FP64 resid",
        "value": best["mdof_solves_per_s"], "unit": "MDOF-solves/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": 1e3 * best["seconds"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-in-f64-IR", "data": "synthetic (seeded lognormal K per instance)",
        "config": {"workload": f"batched ensemble: {count} 2-D flow instances, quad side "
                               f"log-uniform 64-512, {steps} implicit steps each, "
                               "Jacobi-PCG to 1e-10",
                   "dof_total": sum(sc.n for sc in scs)},
        "runs": res,
        "speedup_vs_plain_fp64": best["mdof_solves_per_s"] / res["plain_fp64"]["mdof_solves_per_s"],
        "speedup_vs_group_fp64": best["mdof_solves_per_s"] / res["group_fp64"]["mdof_solves_per_s"],
    }))


def coupled_mode(size, steps):
    """Coupled LIQUID_FLOW -> HEAT_TRANSPORT steps on size^3 cubes split into
    6 tets each: per time step a flow solve (Jacobi-PCG), the Gauss-point
    Darcy velocity of the new pressure written into the heat system on the
    device (k_gauss_velocity), and a heat solve (Jacobi-BiCGStab), both to
    1e-10.  Timed in FP64 and FP32-IR after a one-step warm-up; the velocity
    kernel alone is timed over 50 launches with HIP events; the oracle's
    Cal_Velocity restatement on the same mesh is its CPU baseline."""
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    from oracle import ffi
    flow, heat = W.coupled_flow_heat(size, n_steps=steps)
    warm = [W.Scenario(**{**sc.__dict__, "steps": sc.steps[:1]}) for sc in (flow, heat)]
    dofs = (flow.n + heat.n) * steps
    res = {}
    for name, prec in (("fp64", ox.FP64), ("fp32_ir", ox.FP32_IR)):
        systems = W.make_coupled_systems(flow, heat, prec)
        W.run_coupled_gpu(*warm, precision=prec, systems=systems)
        systems[1].synchronize()
        systems[2].synchronize()
        t0 = time.perf_counter()
        out = W.run_coupled_gpu(flow, heat, precision=prec, systems=systems)
        dt = time.perf_counter() - t0
        infos = out["info_flow"] + out["info_heat"]
        assert all(i["converged"] == 1 and i["error"] <= 1e-10 for i in infos), infos
        res[name] = dict(seconds=dt, mdof_solves_per_s=dofs / dt / 1e6,
                         flow_iterations=[i["iterations"] for i in out["info_flow"]],
                         heat_iterations=[i["iterations"] for i in out["info_heat"]])
        _log(f"coupled {name}: {dt:.2f} s, {dofs / dt / 1e6:.1f} MDOF-solves/s")
        if prec == ox.FP64:
            _, sf, sh = systems
            sf.profile(True)
            for _ in range(50):
                sf.gauss_velocity(sh, host=False)
            ms, launches, bpl = sf.profile_read(ox._lib.PROF_VELOCITY)
            sf.profile(False)
            vel_kernel = dict(kernel="k_gauss_velocity<TET4,iso>", avg_ms=ms / launches,
                              launches=launches, bytes_per_launch=bpl,
                              gbs=bpl / (ms / launches) / 1e6)
            u = sf.get_vector(ox.VEC_X)
            pf = ffi.OracleProblem(flow.mesh, process=flow.process, mass_coef=flow.mass_coef,
                                   cond=flow.cond, mu=flow.mu, rho_g=flow.rho_g)
            t0 = time.perf_counter()
            vr = pf.gauss_velocity(u)
            t_cpu = time.perf_counter() - t0
            vg = sf.gauss_velocity(None, host=True)
            vel_err = float(np.abs(vg - vr).max() / np.abs(vr).max())
        for s in systems[1:]:
            s.close()
    best = res["fp32_ir"] if res["fp32_ir"]["seconds"] < res["fp64"]["seconds"] else res["fp64"]
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if size == 64 and os.path.exists(tf):  # PMC passes of this configuration
        traffic = json.load(open(tf)).get(vel_kernel["kernel"])
    print(json.dumps({
        "metric": "coupled flow+heat MDOF-solves/s to 1e-10",
        "value": best["mdof_solves_per_s"], "unit": "MDOF-solves/s", "n_gpus": 1,
        "steps": steps, "warmup": 1, "ms_per_step": 1e3 * best["seconds"] / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-in-f64-IR" if best is res["fp32_ir"] else "f64",
        "data": "synthetic (seeded lognormal permeability, structured tet mesh)",
        "config": {"workload": f"coupled LIQUID_FLOW -> HEAT_TRANSPORT, {size}^3 x 6 tet4, "
                               f"{steps} steps, Gauss-point velocity on the device",
                   "dof_per_process": flow.n, "elements": flow.mesh.n_elem},
        "roofline": dict(bound="hbm", achieved=vel_kernel["gbs"], peak=HBM_PEAK_GBS,
                         unit="GB/s", frac=vel_kernel["gbs"] / HBM_PEAK_GBS, traffic=traffic,
                         **vel_kernel),
        "cpu_baseline": dict(value=flow.mesh.n_elem / t_cpu / 1e6, unit="M elements/s", cores=1,
                             kind="port",
                             sample=f"oracle Cal_Velocity restatement (serial C) over all "
                                    f"{flow.mesh.n_elem} elements, {t_cpu:.2f} s"),
        "velocity_gpu_melem_per_s": flow.mesh.n_elem / vel_kernel["avg_ms"] / 1e3,
        "velocity_max_rel_diff_vs_oracle": vel_err,
        "runs": res,
    }))


def heat_mode(size, steps, cpu=True):
    """Nonsymmetric heat transport on size^3 cubes split into 6 tets each
    (96^3: 5.3M tets, 912,673 nodes): seeded divergence-free Gauss-point
    velocity with grid Peclet 1-20, 60 C inlet face, Crank-Nicolson, `steps`
    steps of Jacobi-BiCGStab on the reference's row-only Dirichlet system to
    1e-10.  Whole job timed in FP64 and FP32-inside-FP64 refinement after a
    one-step warm-up; then one untimed job with HIP events around every SpMV
    launch (the BiCGStab SpMV kernels), and the tuned format's SpMV alone
    (k_cg_spmv, the same row loop) over 50 launches.  CPU baseline: the
    oracle (serial C) on the 48^3 instance, assembly and 150 BiCGStab
    iterations, scaled per element and per nonzero-iteration."""
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    sc = W.tet_heat(size, n_steps=steps)
    warm = W.Scenario(**{**sc.__dict__, "steps": sc.steps[:1]})
    res = {}
    steps_x = {}
    for name, prec in (("fp64", ox.FP64), ("fp32_ir", ox.FP32_IR)):
        system = W.make_system(sc, prec)
        pat, s = system
        s.set_vector(ox.VEC_IC, sc.u0)
        kw = dict(method=ox.BICGSTAB, dirichlet=ox.DIRICHLET_ROW, precision=prec, system=system,
                  ic_on_device=True, fetch_x=False)
        W.run_gpu(warm, **kw)
        s.synchronize()
        t0 = time.perf_counter()
        out = W.run_gpu(sc, **kw)
        s.synchronize()
        dt = time.perf_counter() - t0
        info = out["info"]
        assert all(i["converged"] == 1 and i["error"] <= 1e-10 for i in info), info
        s.profile(True)
        steps_x[name] = W.run_gpu(sc, keep_steps=True, **kw)["steps"]
        kind = ox._lib.PROF_SPMV64 if prec == ox.FP64 else ox._lib.PROF_SPMV32
        ms, launches, bpl = s.profile_read(kind)
        s.profile(False)
        k_ms, k_bytes = s.bench_spmv(prec == ox.FP32_IR, 50)
        nnz, n = s.nnz, s.n
        res[name] = dict(seconds=dt, ms_per_step=1e3 * dt / steps,
                         mdof_solves_per_s=steps * sc.n / dt / 1e6,
                         iterations=[i["iterations"] for i in info],
                         outer=[i["outer"] for i in info],
                         spmv=dict(avg_ms=k_ms, bytes_per_launch=k_bytes, gbs=k_bytes / k_ms / 1e6,
                                   launches=50, kernel_alone=True),
                         spmv_in_solve=dict(avg_ms=ms / max(launches, 1), launches=launches,
                                            bytes_per_launch=bpl, untimed_job=True),
                         spmv_format=s.spmv_format(prec == ox.FP32_IR))
        _log(f"heat {name}: {dt:.2f} s, {steps * sc.n / dt / 1e6:.1f} MDOF-solves/s, "
             f"iterations {res[name]['iterations'][:4]}...")
        if name == "fp64":
            # every step of both legs against a tight FP64 BiCGStab job (untimed)
            errors, ref_x = error_vs_fp64_ref(sc, system, {}, tol_ref=1e-12, method=ox.BICGSTAB,
                                              dirichlet=ox.DIRICHLET_ROW, with_ref=True)
        s.close()
    for name, xs in steps_x.items():
        errors[name] = [float(np.linalg.norm(x - r) / np.linalg.norm(r)) for x, r in zip(xs, ref_x)]
    best = res["fp32_ir"] if res["fp32_ir"]["seconds"] < res["fp64"]["seconds"] else res["fp64"]
    sp = best["spmv"]
    # algorithmic bytes of one SpMV: every nonzero value read once, x read
    # once, y written once
    width = 4 if best is res["fp32_ir"] else 8
    alg = _spmv_alg_bytes(width, nnz, n, best.get("spmv_format"))
    cpu_base = None
    if cpu:
        from oracle import ffi
        m = 48
        small = W.tet_heat(m, n_steps=1)
        ptr, col, diag = ffi.pattern(small.n, small.mesh.conn)
        eq = ffi.OracleEqs(ptr, col, diag)
        st = small.steps[0]
        prob = ffi.OracleProblem(small.mesh, process=small.process, mass_coef=small.mass_coef,
                                 cond=small.cond, adv_coef=small.adv_coef, vel=small.vel,
                                 dt=st.dt, theta=st.theta, u_old=small.u0)
        t0 = time.perf_counter()
        eq.initialize()
        eq.assemble(prob)
        t_asm = time.perf_counter() - t0
        eq.set_known_x(small.bc_idx, small.bc_val, symmetric=False)
        eq.config(2, 1, 1e-30, 150)
        t0 = time.perf_counter()
        o = eq.solve()
        t_it = (time.perf_counter() - t0) / max(o["iterations"], 1)
        nnz_small = int(ptr[-1])
        per_elem = t_asm / small.mesh.n_elem
        per_nnz_it = t_it / nnz_small
        its = best["iterations"]
        nnz_full = (size + 1) ** 3 + 2 * (3 * (size + 1) ** 2 * size + 3 * (size + 1) * size ** 2
                                          + size ** 3)
        total = steps * per_elem * sc.mesh.n_elem + sum(its) * per_nnz_it * nnz_full
        cpu_base = dict(value=steps * sc.n / total / 1e6, unit="MDOF-solves/s", cores=1, kind="port",
                        sample=(f"oracle (serial C) on the {m}^3 x 6 tet instance: assembly "
                                f"{t_asm:.2f} s for {small.mesh.n_elem} tets, {o['iterations']} "
                                f"BiCGStab iterations at {1e3 * t_it:.1f} ms; scaled per element "
                                f"and per nonzero to the {size}^3 job with the GPU's iteration "
                                f"counts"))
    print(json.dumps({
        "metric": "heat transport MDOF-solves/s to 1e-10",
        "value": best["mdof_solves_per_s"], "unit": "MDOF-solves/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": 1e3 * best["seconds"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-in-f64-IR" if best is res["fp32_ir"] else "f64",
        "data": "synthetic (seeded divergence-free velocity, structured tet mesh)",
        "config": {"workload": f"HEAT_TRANSPORT, {size}^3 x 6 tet4, Crank-Nicolson, {steps} steps, "
                               f"Jacobi-BiCGStab, row-only Dirichlet", "dof": sc.n,
                   "elements": sc.mesh.n_elem},
        "roofline": dict(bound="hbm", achieved=alg / sp["avg_ms"] / 1e6, peak=HBM_PEAK_GBS,
                         unit="GB/s", frac=alg / sp["avg_ms"] / 1e6 / HBM_PEAK_GBS, traffic=None,
                         kernel="k_cg_spmv<float>" if best is res["fp32_ir"] else "k_cg_spmv<double>",
                         avg_launch_ms=sp["avg_ms"], bytes_per_launch=alg,
                         stored_bytes_per_launch=sp["bytes_per_launch"],
                         format=best["spmv_format"]),
        "error_vs_fp64_ref": errors,
        "cpu_baseline": cpu_base,
        "runs": res,
    }))


def quad_mode(steps_timed=5, cpu=True):
    """The CPU-runnable configuration: 2-D transient confined flow on 200 x
    200 quads (40,401 DOF), lognormal K, 10 implicit steps of Jacobi-PCG
    (symmetric Dirichlet) to 1e-10.  The whole 10-step job timed on the GPU
    (FP64 and FP32-IR) `steps_timed` times after a warm-up; the CPU baseline
    is the oracle (serial C) running the identical job end to end (assembly,
    sources, Dirichlet, CG) -- the full workload, no scaling."""
    import ogs5_amd as ox
    from ogs5_amd import workloads as W
    sc = W.quad_flow()
    res = {}
    for name, prec in (("fp64", ox.FP64), ("fp32_ir", ox.FP32_IR)):
        system = W.make_system(sc, prec)
        pat, s = system
        s.set_vector(ox.VEC_IC, sc.u0)
        kw = dict(method=ox.CG, dirichlet=ox.DIRICHLET_SYMMETRIC, precision=prec, system=system,
                  ic_on_device=True, fetch_x=False)
        W.run_gpu(sc, **kw)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps_timed):
            out = W.run_gpu(sc, **kw)
        s.synchronize()
        dt = (time.perf_counter() - t0) / steps_timed
        info = out["info"]
        assert all(i["converged"] == 1 and i["error"] <= 1e-10 for i in info), info
        its = [i["iterations"] for i in info]
        res[name] = dict(seconds_per_job=dt, mdof_solves_per_s=len(sc.steps) * sc.n / dt / 1e6,
                         iterations=its, us_per_iteration=1e6 * dt / max(sum(its), 1),
                         spmv_format=s.spmv_format(prec == ox.FP32_IR))
        _log(f"quad {name}: {1e3 * dt:.1f} ms per job, {len(sc.steps) * sc.n / dt / 1e6:.2f} "
             f"MDOF-solves/s, iterations {its[:3]}...")
        s.close()
    best = res["fp32_ir"] if res["fp32_ir"]["seconds_per_job"] < res["fp64"]["seconds_per_job"] \
        else res["fp64"]
    cpu_base = None
    if cpu:
        from oracle import driver, ffi
        t0 = time.perf_counter()
        ref = driver.run_oracle(sc, ffi, method=5, symmetric=True)
        t_cpu = time.perf_counter() - t0
        x = W.run_gpu(sc, method=ox.CG, dirichlet=ox.DIRICHLET_SYMMETRIC, precision=ox.FP64)["x"]
        rel = float(np.linalg.norm(x - ref["x"]) / np.linalg.norm(ref["x"]))
        cpu_base = dict(value=len(sc.steps) * sc.n / t_cpu / 1e6, unit="MDOF-solves/s", cores=1,
                        kind="port",
                        sample=(f"oracle (serial C) on the whole job: {len(sc.steps)} steps, "
                                f"{sum(i['iterations'] for i in ref['info'])} CG iterations, "
                                f"{t_cpu:.2f} s; final solution vs the GPU's FP64 job: relative "
                                f"L2 {rel:.1e}"))
    print(json.dumps({
        "metric": "MDOF-solves/s to 1e-10 FP64 resid",
        "value": best["mdof_solves_per_s"], "unit": "MDOF-solves/s", "n_gpus": 1,
        "steps": steps_timed, "warmup": 1, "ms_per_step": 1e3 * best["seconds_per_job"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-in-f64-IR" if best is res["fp32_ir"] else "f64",
        "data": "synthetic (seeded lognormal K, structured quad mesh)",
        "config": {"workload": "2-D transient confined flow, 200 x 200 quad4, 10 implicit steps, "
                               "Jacobi-PCG to 1e-10", "dof": sc.n},
        "roofline": None,
        "cpu_baseline": cpu_base,
        "note": "latency-bound (a few microseconds per kernel, three kernels per CG iteration); "
                "the HBM roofline does not apply at this size",
        "runs": res,
    }))


def serve_runs(count, rate, plain=True, fp64=True):
    """Served stream: `count` requests, Poisson arrivals at `rate`/s, 2-D
    quads 32^2-512^2 and 3-D hexes 16^3-96^3, 1-20 steps each.  The
    continuous-batching server in FP32-IR (and FP64), and the plain method
    (one request at a time, solo FP64 solves) on the same stream."""
    from ogs5_amd import _lib as L
    from ogs5_amd import workloads as W
    from ogs5_amd.serve import Server, serve_sequential
    specs = W.request_stream(count, rate)
    res = {}
    runs = [("batched_fp32_ir", L.FP32_IR)] + ([("batched_fp64", L.FP64)] if fp64 else [])
    for name, prec in runs:
        out = Server(prec).serve(specs)
        out.pop("jobs")
        res[name] = out
        _log(f"serve {name}: {out['requests_per_s']:.1f} req/s, "
             f"{out['mdof_solves_per_s']:.1f} MDOF-solves/s, p50 {out['latency_ms']['p50']:.0f} ms, "
             f"p99 {out['latency_ms']['p99']:.0f} ms")
    if plain:
        out = serve_sequential(specs)
        out.pop("jobs")
        res["plain_fp64"] = out
        _log(f"serve plain_fp64: {out['requests_per_s']:.1f} req/s, p50 "
             f"{out['latency_ms']['p50']:.0f} ms, p99 {out['latency_ms']['p99']:.0f} ms")
    return res


def serve_mode(count, rate, count2, rate2):
    res = serve_runs(count, rate)
    # the same kind of stream at a load one GPU sustains: latency is then the
    # service time rather than the queue
    res_sus = serve_runs(count2, rate2, fp64=False) if count2 > 0 else {}
    best = res["batched_fp32_ir"]
    print(json.dumps({
        "metric": "served p50/p99 latency; MDOF-solves/s to 1e-10 FP64 resid",
        "value": best["mdof_solves_per_s"], "unit": "MDOF-solves/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": 1e3 * best["makespan_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32-in-f64-IR", "data": "synthetic request stream (seeded)",
        "config": {"workload": f"served stream: {count} requests, Poisson {rate}/s, quads "
                               "32^2-512^2 and hexes 16^3-96^3, 1-20 steps, Jacobi-PCG to 1e-10"},
        "served_p50_ms": best["latency_ms"]["p50"], "served_p99_ms": best["latency_ms"]["p99"],
        "runs": res,
        "sustainable_load": dict(rate_per_s=rate2, requests=count2, runs=res_sus),
        "speedup_vs_plain_fp64": best["mdof_solves_per_s"] / res["plain_fp64"]["mdof_solves_per_s"],
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["full", "spmv", "ensemble", "serve", "coupled", "heat", "quad",
                                       "assembly"],
                    default="full")
    ap.add_argument("--coupled-size", type=int, default=64)
    ap.add_argument("--heat-size", type=int, default=96)
    ap.add_argument("--heat-steps", type=int, default=20)
    ap.add_argument("--coupled-steps", type=int, default=10)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--ens-steps", type=int, default=10)
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--requests2", type=int, default=150,
                    help="requests of the second served stream (--mode serve), at --rate2")
    ap.add_argument("--rate2", type=float, default=15.0)
    ap.add_argument("--ensemble-instances", type=int, default=256,
                    help="instances of the ensemble comparison in the full mode (0: none)")
    ap.add_argument("--serve-requests", type=int, default=100,
                    help="requests of the short served run in the full mode (0: none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.mode == "spmv":
        spmv_mode(args.size)
        return
    if args.mode == "assembly":
        assembly_mode()
        return
    if args.mode == "ensemble":
        ensemble_mode(args.instances, args.ens_steps)
        return
    if args.mode == "serve":
        serve_mode(args.requests, args.rate, args.requests2, args.rate2)
        return
    if args.mode == "coupled":
        coupled_mode(args.coupled_size, args.coupled_steps)
        return
    if args.mode == "heat":
        heat_mode(args.heat_size, args.heat_steps, cpu=not args.no_cpu)
        return
    if args.mode == "quad":
        quad_mode(cpu=not args.no_cpu)
        return
    import ogs5_amd as ox
    from ogs5_amd import workloads as W

    sc = W.hex_flow(args.size)
    fp64 = gpu_leg(sc, ox.FP64, args.steps, args.warmup)
    ir = gpu_leg(sc, ox.FP32_IR, args.steps, args.warmup)
    # error of every step's solution of both solvers against the job solved
    # tight in FP64 (untimed)
    errors = error_vs_fp64_ref(sc, fp64["system"], {"fp64": fp64.pop("steps_x"),
                                                    "fp32_ir": ir.pop("steps_x")})
    for leg in (ir, fp64):
        leg.pop("system")[1].close()
    best = ir if ir["mdof_solves_per_s"] > fp64["mdof_solves_per_s"] else fp64
    best_name = "fp32_ir" if best is ir else "fp64"
    sp = best["spmv"]
    peak = HBM_PEAK_GBS
    # algorithmic bytes of one SpMV: every value the format must read once
    # (half the off-diagonal ones for the symmetric diagonal copy), x read
    # once, y written once (no index, padding or stored zero is counted; the
    # format's own stored bytes ride along as stored_bytes_per_launch)
    width = 4 if best is ir else 8
    alg = _spmv_alg_bytes(width, _nnz(sc), sc.n, best["spmv_format"])
    achieved = alg / sp["avg_ms"] / 1e6
    in_ms = best["spmv_in_solve"]["avg_ms"]
    roof = dict(bound="hbm", achieved=achieved, peak=peak, unit="GB/s",
                frac=achieved / peak, traffic=None,
                kernel=("k_cg_spmv<float>" if best is ir else "k_cg_spmv<double>"),
                avg_launch_ms=sp["avg_ms"], bytes_per_launch=alg,
                stored_bytes_per_launch=sp["bytes_per_launch"],
                frac_stored=sp["gbs"] / peak if sp["gbs"] else None,
                format=best["spmv_format"],
                # the same kernel inside the solve (events around every launch
                # of one untimed job): p is rewritten between launches
                frac_in_solve=alg / in_ms / 1e6 / peak if in_ms else None,
                avg_launch_ms_in_solve=in_ms)
    # HBM bytes per dispatch of this kernel and format, from the FETCH_SIZE /
    # WRITE_SIZE passes over `bench.py --mode spmv` (profiles/spmv_hex128_traffic.md)
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        key = f"k_cg_spmv<{best['spmv_format']},{'float' if best is ir else 'double'}>"
        tr = json.load(open(tf)).get(key)
        if tr:
            roof["traffic"] = tr
            roof["traffic_source"] = key
    copy_gbs = ox._lib.stream_copy_gbs(1 << 30, 10)
    roof["copy_kernel_gbs"] = copy_gbs
    out = {
        "metric": "MDOF-solves/s to 1e-10 FP64 resid; SpMV HBM GB/s; served p50/p99 latency",
        "value": best["mdof_solves_per_s"],
        "unit": "MDOF-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best is fp64 else "f32-in-f64-IR",
        "data": "synthetic (seeded lognormal K, structured hex mesh)",
        "config": {"workload": f"3D liquid flow, {args.size}^3 hex8, steady + 5 transient "
                               "steps, Jacobi-PCG to 1e-10",
                   "dof": sc.n, "nnz": _nnz(sc), "solver": best_name},
        "roofline": roof,
        "cpu_baseline": None,
        "error_vs_fp64_ref": errors,
        "fp64": fp64,
        "fp32_ir": ir,
        "ir_speedup_vs_fp64": ir["mdof_solves_per_s"] / fp64["mdof_solves_per_s"],
        "spmv_fp64_hbm_gbs": fp64["spmv"]["gbs"],
        "served_p50_ms": None,
        "served_p99_ms": None,
    }
    if args.serve_requests > 0:
        sv = serve_runs(args.serve_requests, args.rate, plain=False, fp64=False)["batched_fp32_ir"]
        out["served_p50_ms"] = sv["latency_ms"]["p50"]
        out["served_p99_ms"] = sv["latency_ms"]["p99"]
        out["served"] = dict(requests=sv["requests"], rate_per_s=args.rate,
                             requests_per_s=sv["requests_per_s"],
                             mdof_solves_per_s=sv["mdof_solves_per_s"],
                             note="short served run (continuous batching, FP32-IR); "
                                  "bench.py --mode serve runs the full comparison")
    if args.ensemble_instances > 0:
        # the batched-ensemble configuration: the FP32-IR group against the
        # plain FP64 solver (one instance at a time), both to 1e-10
        runs, scs = ensemble_runs(args.ensemble_instances, args.ens_steps, group_fp64=False)
        out["ensemble"] = dict(
            workload=f"{args.ensemble_instances} 2-D flow instances, quad side log-uniform "
                     f"64-512, {args.ens_steps} implicit steps each, Jacobi-PCG to 1e-10",
            dof_total=sum(x.n for x in scs), runs=runs,
            speedup_vs_plain_fp64=runs["group_fp32_ir"]["mdof_solves_per_s"] /
            runs["plain_fp64"]["mdof_solves_per_s"])
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, fp64["iterations"])
    print(json.dumps(out))


def _nnz(sc):
    n1 = sc.mesh.shape[0] + 1
    return (3 * n1 - 2) ** 3


def _spmv_alg_bytes(width, nnz, n, fmt):
    """Algorithmic bytes of one SpMV: every stored value the format must read
    once (all nnz; for the symmetric diagonal copy, format 25, the diagonal
    and one triangle: (nnz + n) / 2), x read once, y written once."""
    vals = (nnz + n) // 2 if fmt == 25 else nnz
    return width * vals + 2 * width * n


if __name__ == "__main__":
    main()
