"""Serving loop: solve requests of different mesh sizes as they arrive,
with continuous batching on one GPU.

Every admitted request owns a device-resident system (matrix, Krylov
vectors, element data) for its whole life; all running solves share one
ogsx group, whose Krylov phases are launched once over every running
request.  A request that arrives while others are on the GPU is admitted
as soon as a slot and its HBM footprint (ogsx_pattern_footprint) fit, and
its solves join the running batch between two iteration batches.  Host
work of a request (mesh, K field, CRS pattern, colouring and device
layouts) runs on a worker pool, and its device set-up and tear-down on one
more thread, off the GPU loop.

Each step of a request is OGS-5's time step on its own system: Initialize
-> assembly -> source terms -> Dirichlet -> solve -> copy-back
(CRFProcess::Execute, FEM/rf_pcs.cpp:4695-4853).  The steps after a
request's first are prepared for all the requests that start one together
in one batched set of launches (ogsx_group_step), with each request's A and
b bit-identical to its own calls; the group solve returns the same iterates
as a solo solve, so every result is bit-identical to the request solved
alone.

Latency of a request = completion - arrival (queueing, host setup and all
its steps).
"""
from __future__ import annotations

import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib as L
from . import workloads as W


@dataclass
class Job:
    spec: W.RequestSpec
    sc: W.Scenario
    pattern: L.Pattern
    footprint: int
    t_ready: float
    system: Optional[L.System] = None
    step: int = 0
    infos: list = field(default_factory=list)
    t_admit: float = 0.0
    t_done: float = 0.0
    x: Optional[np.ndarray] = None


class MeshCache:
    """Meshes and CRS patterns shared by requests of the same kind and size
    (both depend on nothing else): built once, by the first request that
    needs them, and kept for the most recent `size` kinds/sizes.  Patterns
    are read-only once built, so concurrent systems share one."""

    def __init__(self, size: int = 32):
        self.size = size
        self.lock = threading.Lock()
        self.entries: "OrderedDict[tuple, Future]" = OrderedDict()

    def get(self, kind: str, side: int):
        key = (kind, side)
        with self.lock:
            fut = self.entries.get(key)
            mine = fut is None
            if mine:
                fut = Future()
                self.entries[key] = fut
                while len(self.entries) > self.size:
                    self.entries.popitem(last=False)  # dropped patterns free with their last user
            self.entries.move_to_end(key)
        if mine:
            try:
                mesh = W.request_mesh(kind, side)
                fut.set_result((mesh, L.Pattern(mesh.n_nodes, mesh.etype, mesh.conn)))
            except BaseException as e:
                fut.set_exception(e)
        return fut.result()


def _setup(spec: W.RequestSpec, precision: int, t0: float, cache: MeshCache) -> Job:
    """Host-side work of a request (runs on the worker pool; no GPU calls)."""
    mesh, pat = cache.get(spec.kind, spec.side)
    sc = W.request_scenario(spec, mesh)
    return Job(spec, sc, pat, pat.footprint(precision), time.perf_counter() - t0)


def _materialize(job: Job, precision: int) -> Job:
    """Device set-up of an admitted request: system, element data, initial
    condition (on the request's own stream, off the GPU loop's thread)."""
    s = L.System(job.pattern, precision)
    s.set_elements(**job.sc.element_kwargs())
    s.set_vector(L.VEC_UOLD, job.sc.u0)
    s.set_vector(L.VEC_X, job.sc.u0)
    s.synchronize()
    job.system = s
    return job


def _teardown(job: Job):
    job.system.close()
    job.system = None
    job.pattern = None  # shared through the mesh cache


class Server:
    """Continuous-batching solver of flow requests on one GPU.

    precision     OGSX_FP32_IR (FP32 Krylov inside FP64 refinement) or FP64
    hbm_budget    device bytes the resident requests may hold together
    max_active    requests resident at once
    batch_iters   Krylov iterations between two admission / completion checks
    """

    def __init__(self, precision: int = L.FP32_IR, hbm_budget: int = 64 << 30,
                 max_active: int = 64, workers: int = 8, batch_iters: int = 64,
                 use_graphs: bool = True, tol: float = 1e-10, max_iter: int = 100000):
        self.precision = precision
        self.hbm_budget = int(hbm_budget)
        self.max_active = max_active
        self.workers = workers
        self.batch_iters = batch_iters
        self.use_graphs = use_graphs
        self.tol = tol
        self.max_iter = max_iter

    def _begin_steps(self, group, jobs: List[Job]):
        """Prepares each job's next step and starts its solve: a first step
        (which uploads the request's source and Dirichlet lists) or a restart
        from the initial condition on the job's own stream, every other step
        of all the jobs in one batched preparation (Group.step, u_old <- x
        included); one group call starts the solves."""
        if not jobs:
            return
        # a request may carry a tighter iteration limit ($LINEAR_SOLVER
        # max_iterations of its own deck); one start call per limit
        limits = {}
        for j in jobs:
            limits.setdefault(min(self.max_iter, j.sc.max_iter), []).append(j)
        for max_iter, grp in limits.items():
            many = [j for j in grp if W.batchable(j.sc, j.step)]
            one = [j for j in grp if not W.batchable(j.sc, j.step)]
            for j in one:
                if j.step > 0:
                    j.system.copy_vector(L.VEC_UOLD, L.VEC_X)
                W._prepare_step(j.sc, j.system, j.step, L.DIRICHLET_SYMMETRIC)
            if one:
                group.start([j.system for j in one], method=L.CG, tol=self.tol,
                            max_iter=max_iter)
            if many:
                group.step([j.system for j in many],
                           [W.step_params(j.sc, j.step, L.DIRICHLET_SYMMETRIC) for j in many],
                           method=L.CG, tol=self.tol, max_iter=max_iter)

    def serve(self, specs: List[W.RequestSpec], keep_x: bool = False,
              log=None) -> dict:
        """Serves the requests (arrival times relative to the call) and
        returns per-request records and the aggregate figures."""
        L.device_count()  # the runtime initialises on this thread
        group = L.Group(self.precision, self.use_graphs)
        order = sorted(specs, key=lambda s: s.arrival)
        pool = ThreadPoolExecutor(self.workers)
        # creates admitted requests' systems and destroys finished ones while
        # the loop keeps the group running
        devq = ThreadPoolExecutor(1)
        cache = MeshCache()
        t0 = time.perf_counter()
        setups = deque()  # futures in arrival order
        ready: deque = deque()
        creating: deque = deque()  # futures of admitted requests being set up
        live: dict = {}
        done: List[Job] = []
        used = 0
        peak_used = 0
        peak_live = 0
        nxt = 0
        batches = 0
        try:
            while len(done) < len(order):
                now = time.perf_counter() - t0
                while nxt < len(order) and order[nxt].arrival <= now:
                    setups.append(pool.submit(_setup, order[nxt], self.precision, t0, cache))
                    nxt += 1
                while setups and setups[0].done():
                    ready.append(setups.popleft().result())
                # admission in arrival order, by slots and HBM footprint
                while ready and len(live) + len(creating) < self.max_active and \
                        used + ready[0].footprint <= self.hbm_budget:
                    job = ready.popleft()
                    used += job.footprint
                    creating.append(devq.submit(_materialize, job, self.precision))
                if ready and not live and not creating and ready[0].footprint > self.hbm_budget:
                    raise L.OgsxError(f"request {ready[0].spec.rid} needs "
                                      f"{ready[0].footprint} bytes, above the HBM budget")
                joining = []
                while creating and creating[0].done():
                    job = creating.popleft().result()
                    group.add(job.system)
                    job.t_admit = time.perf_counter() - t0
                    live[job.spec.rid] = job
                    joining.append(job)
                self._begin_steps(group, joining)
                peak_used = max(peak_used, used)
                peak_live = max(peak_live, len(live))
                if not live:
                    if creating:
                        creating[0].result()
                    elif nxt < len(order) and not setups and not ready:
                        time.sleep(max(0.0, min(order[nxt].arrival - now, 0.002)))
                    else:
                        time.sleep(0.0005)
                    continue
                group.advance(self.batch_iters)
                batches += 1
                stepping = []
                for rid in list(live):
                    job = live[rid]
                    active, info = L.solve_status(job.system)
                    if active:
                        continue
                    job.infos.append(info)
                    s = job.system
                    job.step += 1
                    if job.step < job.spec.steps:
                        stepping.append(job)
                        continue
                    if keep_x:
                        job.x = s.get_vector(L.VEC_X)
                    else:
                        s.synchronize()
                    job.t_done = time.perf_counter() - t0
                    group.remove(s)
                    devq.submit(_teardown, job)
                    used -= job.footprint
                    del live[rid]
                    done.append(job)
                    if log:
                        log(job)
                self._begin_steps(group, stepping)
        finally:
            pool.shutdown(wait=True)
            devq.shutdown(wait=True)
            group.close()
        return self._summary(done, time.perf_counter() - t0, peak_used, peak_live, batches)

    def _summary(self, done: List[Job], wall: float, peak_used: int, peak_live: int,
                 batches: int) -> dict:
        lat = np.array([j.t_done - j.spec.arrival for j in done])
        wait = np.array([j.t_admit - j.spec.arrival for j in done])
        dof_solves = sum(j.sc.n * len(j.infos) for j in done)
        first = min(j.spec.arrival for j in done)
        last = max(j.t_done for j in done)
        bad = [j.spec.rid for j in done if any(i["converged"] != 1 or i["error"] > self.tol
                                              for i in j.infos)]
        return dict(
            requests=len(done), wall_s=wall, makespan_s=last - first,
            mdof_solves_per_s=dof_solves / (last - first) / 1e6,
            requests_per_s=len(done) / (last - first),
            latency_ms=dict(p50=1e3 * float(np.percentile(lat, 50)),
                            p99=1e3 * float(np.percentile(lat, 99)),
                            mean=1e3 * float(lat.mean()), max=1e3 * float(lat.max())),
            admission_wait_ms=dict(p50=1e3 * float(np.percentile(wait, 50)),
                                   p99=1e3 * float(np.percentile(wait, 99))),
            peak_resident_bytes=peak_used, peak_resident_requests=peak_live,
            batches=batches, not_converged=bad,
            jobs=done,
        )


def serve_sequential(specs: List[W.RequestSpec], precision: int = L.FP64,
                     tol: float = 1e-10, max_iter: int = 100000, workers: int = 8) -> dict:
    """The plain method on the same stream: one request at a time, each
    step a solo ogsx_solve (host setup still overlaps on the pool)."""
    L.device_count()
    order = sorted(specs, key=lambda s: s.arrival)
    pool = ThreadPoolExecutor(workers)
    cache = MeshCache()
    t0 = time.perf_counter()
    futs = {}
    done = []
    nxt = 0
    try:
        for spec in order:
            now = time.perf_counter() - t0
            if spec.arrival > now:
                time.sleep(spec.arrival - now)
            # setups of everything that has arrived are queued on the pool
            now = time.perf_counter() - t0
            while nxt < len(order) and order[nxt].arrival <= now:
                futs[order[nxt].rid] = pool.submit(_setup, order[nxt], precision, t0, cache)
                nxt += 1
            if spec.rid not in futs:
                futs[spec.rid] = pool.submit(_setup, spec, precision, t0, cache)
                nxt = max(nxt, order.index(spec) + 1)
            job = futs.pop(spec.rid).result()
            s = L.System(job.pattern, precision)
            s.set_elements(**job.sc.element_kwargs())
            job.system = s
            job.t_admit = time.perf_counter() - t0
            res = W.run_gpu(job.sc, precision=precision, system=(job.pattern, s), tol=tol)
            job.infos = res["info"]
            job.t_done = time.perf_counter() - t0
            s.close()
            job.pattern = None
            job.system = None
            done.append(job)
    finally:
        pool.shutdown(wait=True)
    srv = Server(precision, tol=tol)
    return srv._summary(done, time.perf_counter() - t0, 0, 1, 0)
