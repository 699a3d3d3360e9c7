"""Serving loop for S-parameter sweep jobs on one device.

Jobs (a mesh with modal ports, swept over all its (port, mode) channels)
arrive over time.  The loop keeps every admitted job resident on the device
— fields, port states, DTBC histories, V/I records — and advances all of
them together: each resident job owns a stream and a captured graph of its
steps, and the host enqueues one chunk of steps per resident job per round,
so the jobs' kernels overlap on the device.  A job that arrives while others
run is admitted as soon as a slot and its HBM footprint are free (first fit
in arrival order); a finished job's S-matrix is assembled on the host and
its memory returned.  ``mode="sequential"`` runs the same jobs one at a
time, in arrival order, for comparison.

The reference has no serving tier; its unit of work is one
``AnalysisScatteringTD.run`` (src/magnelio/analysis/scattering_td.py:1252-1330),
which a job here reproduces as one batched sweep (``sweep.ScatteringSweep``).
Results do not depend on the mix: every job has its own buffers and every
kernel is deterministic, so a job's S-matrix in the mix equals its S-matrix
alone bit for bit.  A job may ask for a float64 check of its result
(``SweepJob.fp64_check``): the check solver is admitted with the job (its
footprint counted), advanced on the job's stream and its S-parameter error
recorded with the job.  Small jobs (``pack_cells``) step together: one
packed group (``pack.PackedGroup``) advances all of them per step in shared
launches on a stream of its own, rebuilt whenever a member joins or leaves.
Host preparation (material matrices, coefficients)
runs for at most twice ``max_active`` waiting jobs at a time, and admitted
solvers drop their float64 host copies once on the device, so host memory
stays bounded however far arrivals outrun service.
"""

from __future__ import annotations

import heapq
import sys
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from . import kernels, pack
from .sweep import SParameterResult, ScatteringSweep


@dataclass
class BuildPlan:
    """A job's host build split into tasks that a pool can run apart.

    ``tasks``: {name: (cache_key or None, callable)}; tasks with equal
    non-None keys are run once per server and their result shared (a port
    model depends only on its feed cross-section, not on the job).
    ``assemble({name: result})`` returns (mesh, ports, dt).  The callables
    must be picklable for a process pool."""

    tasks: dict
    assemble: object

    def __call__(self):
        return self.assemble({name: fn() for name, (_, fn) in self.tasks.items()})


@dataclass
class SweepJob:
    """One sweep request.  ``build()`` (a callable or a :class:`BuildPlan`)
    returns (mesh, ports, dt) on the host (mode solving included); it runs in
    worker threads or processes before admission."""

    job_id: int
    arrival: float  # seconds after the start of the run
    build: object
    n_steps: int
    f_axis: np.ndarray
    f_max: float
    precision: str = "single"
    waveforms: dict | None = None
    # optional: CPML faces (default: PEC walls on the non-port faces) and a
    # factory of DFT monitors recorded in every field set
    boundary_conditions: dict | None = None
    monitors_factory: object = None
    # float64 check of the result ("off", "column", "all"; ScatteringSweep):
    # its solver is admitted and stepped with the job, its footprint counted
    fp64_check: str = "off"


def footprint_bytes(dims, n_channels: int, K: int, n_steps: int, precision: str = "single",
                    psi_words: int = 0, dft_bytes: int = 0) -> int:
    """Device bytes of a resident sweep: K field sets plus the shared alpha/beta
    and masses, the DTBC histories (w, s) and V/I records of every channel of
    every set in float64, the port work areas, the CPML psi slabs (psi_words
    per set, plus their one-byte air masks) and the DFT monitors' bins and
    sample buffers (dft_bytes, all sets)."""
    nx, ny, nz = dims
    ne, nh = kernels.n_e(nx, ny, nz), kernels.n_h(nx, ny, nz)
    w = 4 if precision == "single" else 8
    fields = w * (ne + nh) * K
    coeffs = w * (ne + nh) * 3  # alpha, beta and the mass vector of each side
    per_channel = 8 * 4 * max(n_steps, 1)  # w_hist, s_hist, V_rec, I_rec
    psi = (w * K + 1) * int(psi_words)
    return int(fields + coeffs + per_channel * n_channels * K + psi + int(dft_bytes) + (64 << 20))


def job_extras(boundary_conditions, monitors, grid, K: int) -> tuple[int, int]:
    """(psi words per set, DFT monitor bytes) of a job's CPML faces and monitors."""
    psi = 0
    for bc in (boundary_conditions or {}).values():
        if getattr(bc, "kind", "") == "CPML":
            e_sh, h_sh = bc.psi_shapes(grid.dims)
            psi += sum(int(np.prod(sh)) for sh in e_sh + h_sh)
    dft = sum(int(m.footprint(grid, boundary_conditions, K)) for m in (monitors or []))
    return psi, dft


@dataclass
class JobRecord:
    job_id: int
    arrival: float
    prepared: float = 0.0  # host preparation done (built and prepared)
    admitted: float = 0.0
    finished: float = 0.0
    cells: int = 0
    K: int = 0
    n_steps: int = 0
    footprint: int = 0
    result: SParameterResult | None = None
    fp64: dict | None = None  # the float64 check of the result, when asked for
    check_sets: int = 0  # field sets of the float64 check solver
    packed: bool = False  # stepped in the packed group of small jobs
    status: str = "pending"  # "done", "rejected" (footprint over the budget) or "cancelled"
    reason: str = ""

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def cell_updates(self) -> float:
        return float(self.cells) * self.K * self.n_steps


@dataclass
class ServeReport:
    mode: str
    records: list  # every job, in arrival order (see JobRecord.status)
    wall: float
    peak_resident: int = 0
    peak_bytes: int = 0

    @property
    def completed(self) -> list:
        """Records of the jobs that ran to their S-matrix; the statistics
        below are over these."""
        return [r for r in self.records if r.status == "done"]

    def latencies(self) -> np.ndarray:
        return np.array([r.latency for r in self.completed])

    def breakdown(self) -> dict:
        """Median seconds per phase: host preparation, wait for admission,
        device run (admission to S-matrix)."""
        done = self.completed
        prep = np.array([r.prepared - r.arrival for r in done])
        wait = np.array([r.admitted - r.prepared for r in done])
        run = np.array([r.finished - r.admitted for r in done])
        return {"prep_p50": float(np.median(prep)), "wait_p50": float(np.median(wait)),
                "run_p50": float(np.median(run)), "run_p99": float(np.percentile(run, 99))}

    @property
    def p50(self) -> float:
        return float(np.percentile(self.latencies(), 50))

    @property
    def p99(self) -> float:
        return float(np.percentile(self.latencies(), 99))

    @property
    def jobs_per_s(self) -> float:
        return len(self.completed) / self.wall

    @property
    def gcell_per_s(self) -> float:
        return sum(r.cell_updates for r in self.completed) / self.wall / 1e9


@dataclass
class _Active:
    rec: JobRecord
    sweep: ScatteringSweep
    solver: object
    stream: torch.cuda.Stream
    check: object = None  # float64 check solver, stepped after the job's solver
    packed: bool = False  # stepped in the shared packed group, not on its own stream
    events: deque = field(default_factory=deque)

    def completed(self) -> bool:
        n = self.rec.n_steps
        return self.solver.n_completed >= n and (self.check is None or self.check.n_completed >= n)


class SweepServer:
    """Continuous-batching server (see module docstring).

    ``hbm_budget``: bytes the resident jobs may hold (default 85 % of the
    device memory free at start); ``max_active``: resident jobs at once;
    ``chunk``: steps enqueued per job per round; ``inflight``: chunks a job
    may have queued ahead of the device.
    """

    def __init__(self, device: str = "cuda:0", hbm_budget: int | None = None, max_active: int = 8,
                 chunk: int = 256, inflight: int = 2, build_workers: int = 2,
                 use_graphs: bool = True, build_pool=None, progress: int = 0,
                 pack_cells: int = 0):
        self.device = torch.device(device)
        self.hbm_budget = hbm_budget
        self.max_active = int(max_active)
        self.chunk = int(chunk)
        self.inflight = int(inflight)
        self.build_workers = int(build_workers)
        self.use_graphs = use_graphs
        # an executor for job.build (e.g. a process pool: the builds are CPU
        # bound and hold the GIL); None = a thread pool of build_workers
        self.build_pool = build_pool
        self.progress = int(progress)  # report every `progress` finished jobs (0: never)
        # resident jobs of at most this many cells (and packable: no CPML,
        # sources or monitors, no float64 check) step together in shared
        # launches (pack.PackedGroup); 0: every job on its own
        self.pack_cells = int(pack_cells)
        self._cancel: set = set()

    def cancel(self, job_id: int) -> None:
        """Cancel a job of the running ``run`` (or, called before it, of the
        next one) — from any thread.
        A job not yet arrived or built is dropped, a waiting one leaves the
        queue, a resident one is retired once its queued steps drain (its
        memory returns to the budget); it gets no S-matrix and the status
        "cancelled".  A job already done is left alone."""
        self._cancel.add(int(job_id))

    # -- helpers ---------------------------------------------------------
    def _budget(self) -> int:
        if self.hbm_budget is not None:
            return int(self.hbm_budget)
        free, _ = torch.cuda.mem_get_info(self.device)
        return int(0.85 * free)

    def _prepare(self, job: SweepJob, rec: JobRecord, t0: float, built=None):
        """Host half of a job: mesh, ports, material matrices and
        coefficients on the host (no device access).  ``built``: the result
        of ``job.build()`` when it already ran elsewhere."""
        mesh, ports, dt = built if built is not None else job.build()
        sw = ScatteringSweep(mesh, ports, dt=dt, n_steps=job.n_steps, f_axis=job.f_axis,
                             f_max=job.f_max, waveforms=job.waveforms, precision=job.precision,
                             use_graphs=self.use_graphs, device=str(self.device),
                             boundary_conditions=job.boundary_conditions,
                             monitors_factory=job.monitors_factory, fp64_check=job.fp64_check)
        solver = sw.make_solver(sw.excited)
        solver.prepare_host()
        # the float64 check solver takes its host half at admission
        check = sw.make_check_solver()
        rec.cells = int(np.prod(mesh.dims))
        rec.K = len(sw.excited)
        rec.n_steps = job.n_steps
        psi, dft = job_extras(sw.boundary_conditions, solver.monitors, mesh.grid, rec.K)
        rec.footprint = footprint_bytes(mesh.dims, len(sw.channels), rec.K, job.n_steps, job.precision,
                                        psi_words=psi, dft_bytes=dft)
        if check is not None:
            rec.check_sets = len(sw.check_channels())
            psi64, _ = job_extras(sw.boundary_conditions, [], mesh.grid, rec.check_sets)
            rec.footprint += footprint_bytes(mesh.dims, len(sw.channels), rec.check_sets,
                                             job.n_steps, "double", psi_words=psi64)
        rec.prepared = time.perf_counter() - t0
        return sw, solver, check

    @staticmethod
    def _submit_build(executor, build, shared: dict) -> dict:
        """Futures of a job's build tasks; keyed tasks are shared server-wide."""
        if not isinstance(build, BuildPlan):
            return {None: executor.submit(build)}
        futs = {}
        for name, (key, fn) in build.tasks.items():
            if key is None:
                futs[name] = executor.submit(fn)
            else:
                if key not in shared:
                    shared[key] = executor.submit(fn)
                futs[name] = shared[key]
        return futs

    @staticmethod
    def _assemble(build, futs: dict):
        if not isinstance(build, BuildPlan):
            return futs[None].result()
        return build.assemble({name: f.result() for name, f in futs.items()})

    @staticmethod
    def _finish(a: _Active) -> SParameterResult:
        cols = [a.sweep.s_column(a.solver.recorded(k), ch) for k, ch in enumerate(a.sweep.excited)]
        res = SParameterResult.merge(cols)
        if a.check is not None:
            a.rec.fp64 = a.sweep.fp64_error(res, a.check)
        return res

    # -- main loop -------------------------------------------------------
    def run(self, jobs: list[SweepJob], mode: str = "continuous") -> ServeReport:
        if mode not in ("continuous", "sequential"):
            raise ValueError("mode must be 'continuous' or 'sequential'")
        max_active = self.max_active if mode == "continuous" else 1
        budget = self._budget()
        order = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
        recs = {j.job_id: JobRecord(j.job_id, j.arrival) for j in order}
        pool = ThreadPoolExecutor(max_workers=max(1, self.build_workers))
        external = self.build_pool
        futures = {}
        shared: dict = {}  # cache key -> future of a shared build task
        arrivals = [(j.arrival, j.job_id, j) for j in order]
        heapq.heapify(arrivals)
        waiting: list[tuple] = []  # (job, sweep, solver prepared on the host)
        active: list[_Active] = []
        pack_stream = torch.cuda.Stream(device=self.device)
        pack_events: deque = deque()
        pack_key, group = None, None
        used = 0
        peak_resident = peak_bytes = 0
        done = 0
        t0 = time.perf_counter()
        try:
            while done < len(order):
                now = time.perf_counter() - t0
                # cancellations: jobs not yet resident leave at once
                for jid in list(self._cancel):
                    rec = recs.get(jid)
                    if rec is None or rec.status != "pending":
                        continue
                    if any(a.rec is rec for a in active):
                        continue  # retired below, once its queued steps drain
                    futures.pop(jid, None)
                    waiting[:] = [w for w in waiting if w[0].job_id != jid]
                    rec.status, rec.reason = "cancelled", "cancelled before admission"
                    done += 1
                # arrivals start their host-side build (mesh, port modes)
                while arrivals and arrivals[0][0] <= now:
                    _, jid, job = heapq.heappop(arrivals)
                    if recs[jid].status != "pending":
                        continue
                    futures[jid] = (job, self._submit_build(external or pool, job.build, shared))
                # built jobs are prepared (material matrices, coefficients:
                # gigabytes of host memory for a large job) and join the
                # admission queue in arrival order, while fewer than
                # 2 max_active prepared jobs wait
                for jid in sorted(futures, key=lambda i: (recs[i].arrival, i)):
                    if len(waiting) >= 2 * max_active:
                        break
                    job, fut = futures[jid]
                    if all(f.done() for f in fut.values()):
                        built = self._assemble(job.build, fut)
                        waiting.append((job, *self._prepare(job, recs[jid], t0, built)))
                        del futures[jid]
                waiting.sort(key=lambda x: (x[0].arrival, x[0].job_id))
                # admission: first fit in arrival order
                for item in list(waiting):
                    job, sw, solver, check = item
                    rec = recs[job.job_id]
                    if len(active) >= max_active:
                        break
                    if rec.footprint > budget:  # can never fit: rejected, the others go on
                        rec.status = "rejected"
                        rec.reason = (f"needs {rec.footprint / 1e9:.1f} GB, more than the "
                                      f"{budget / 1e9:.1f} GB budget")
                        waiting.remove(item)
                        done += 1
                        continue
                    if used + rec.footprint > budget:
                        if mode == "sequential":
                            break
                        continue
                    # buffers come from the current stream's pool, so a retired
                    # job's memory is reused by the next one; the job's stream
                    # waits for the set-up copies before its first step
                    solver.setup()
                    solver.release_host()
                    if check is not None:
                        check.setup()
                        check.release_host()
                    stream = torch.cuda.Stream(device=self.device)
                    stream.wait_stream(torch.cuda.current_stream(self.device))
                    rec.admitted = time.perf_counter() - t0
                    packed = (self.pack_cells > 0 and rec.cells <= self.pack_cells
                              and check is None and pack.packable(solver))
                    rec.packed = packed
                    active.append(_Active(rec, sw, solver, stream, check, packed))
                    used += rec.footprint
                    waiting.remove(item)
                peak_resident = max(peak_resident, len(active))
                peak_bytes = max(peak_bytes, used)
                # one chunk per resident job per round, up to `inflight` queued
                progressed = False
                # small jobs: one packed group, rebuilt when its members change
                members = [a for a in active if a.packed and not a.completed()
                           and a.rec.job_id not in self._cancel][:pack.MAX_JOBS]
                while pack_events and pack_events[0].query():
                    pack_events.popleft()
                if members and len(pack_events) < self.inflight:
                    key = tuple(a.rec.job_id for a in members)
                    if key != pack_key:
                        group = pack.PackedGroup([a.solver for a in members],
                                                 use_graphs=self.use_graphs)
                        pack_key = key
                        pack_stream.wait_stream(torch.cuda.current_stream(self.device))
                    with torch.cuda.stream(pack_stream):
                        group.launch(self.chunk)
                        ev = torch.cuda.Event()
                        ev.record(pack_stream)
                    pack_events.append(ev)
                    progressed = True
                for a in active:
                    if a.packed:
                        continue
                    while a.events and a.events[0].query():
                        a.events.popleft()
                    if (len(a.events) < self.inflight and not a.completed()
                            and a.rec.job_id not in self._cancel):
                        with torch.cuda.stream(a.stream):
                            if a.solver.n_completed < a.rec.n_steps:
                                a.solver.launch(self.chunk)
                            if a.check is not None and a.check.n_completed < a.rec.n_steps:
                                a.check.launch(self.chunk)
                            ev = torch.cuda.Event()
                            ev.record(a.stream)
                        a.events.append(ev)
                        progressed = True
                # retire finished (or cancelled) jobs
                for a in list(active):
                    cancelled = a.rec.job_id in self._cancel
                    drained = (not pack_events or pack_events[-1].query()) if a.packed else \
                        all(e.query() for e in a.events)
                    if (cancelled or a.completed()) and drained:
                        (pack_stream if a.packed else a.stream).synchronize()
                        if cancelled:
                            a.rec.status, a.rec.reason = "cancelled", "cancelled while resident"
                        else:
                            a.rec.result = self._finish(a)
                            a.rec.status = "done"
                        a.rec.finished = time.perf_counter() - t0
                        used -= a.rec.footprint
                        active.remove(a)
                        del a
                        done += 1
                        progressed = True
                        if self.progress and done % self.progress == 0:
                            print(f"serve[{mode}]: {done}/{len(order)} jobs done at "
                                  f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
                if not progressed:
                    time.sleep(2e-4)
        finally:
            pool.shutdown(wait=True)
            self._cancel.clear()
        wall = time.perf_counter() - t0 - min(j.arrival for j in order)
        return ServeReport(mode, [recs[j.job_id] for j in order], wall, peak_resident, peak_bytes)
