"""GPU serving loop: continuous batching of sweep jobs.

Small seeded jobs of the serving workload (WR-90 guides and crosses of two
with dielectric bricks/spheres, 1-4 modal ports x 1-3 modes) arrive close
together:

* in the mix (several jobs resident, their chunks overlapping on separate
  streams) every job's S-matrix equals the same job run alone, bit for bit;
* with an HBM budget that holds only part of the jobs, admission keeps the
  resident footprint within the budget and every job still completes with
  the same S-matrix;
* the one-job-at-a-time mode gives the same S-matrices;
* latency percentiles and throughput are reported;
* the admission footprint of a job with a CPML face and a DFT monitor
  covers what its solver allocates on the device (psi slabs, bins and sample
  buffers included) within the 64 MiB slack plus 10 %;
* a job whose footprint exceeds the budget is rejected (recorded, no
  S-matrix) while the others complete unchanged; a job cancelled before the
  run and one cancelled from another thread while the loop runs end as
  "cancelled", and the rest still equal their runs alone;
* with packing (jobs up to 40^3 cells share launches, the group rebuilt as
  jobs come and go) every job's S-matrix still equals its run alone;
* jobs that ask for the float64 check carry its error (every S entry of the
  checked column within 1e-4 of the float64 run) and the same S-matrices as
  without it, the check solver stepped in the mix.
"""

import numpy as np
import pytest

import bench
from magnelio_amd.serve import SweepServer
from magnelio_amd.sweep import ScatteringSweep

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jobs():
    return bench.c5_jobs(n_jobs=5, seed=11, rate=20.0, cells=(16 ** 3, 40 ** 3), steps=(300, 700))


@pytest.fixture(scope="module")
def alone(jobs, gpu):
    out = {}
    for j in jobs:
        mesh, ports, dt = j.build()
        print(f"job {j.job_id} alone: dims {mesh.dims}, {sum(p.n_modes for p in ports)} channels, "
              f"{j.n_steps} steps", flush=True)
        sw = ScatteringSweep(mesh, ports, dt=dt, n_steps=j.n_steps, f_axis=j.f_axis, f_max=j.f_max)
        out[j.job_id] = sw.run().s_params
    return out


def _same(a, b):
    return (a.channels == b.channels and a.excitations == b.excitations
            and np.array_equal(a.matrix, b.matrix, equal_nan=True))


def test_mix_equals_alone(jobs, alone, gpu):
    print("continuous mix", flush=True)
    rep = SweepServer(max_active=4, chunk=64).run(jobs, mode="continuous")
    assert rep.peak_resident >= 2
    for r in rep.records:
        assert _same(r.result, alone[r.job_id]), r.job_id
    lat = rep.latencies()
    assert (lat > 0).all() and rep.p99 >= rep.p50 > 0 and rep.jobs_per_s > 0
    print(f"p50 {rep.p50:.3f} s  p99 {rep.p99:.3f} s  {rep.jobs_per_s:.2f} jobs/s  "
          f"peak resident {rep.peak_resident}")


def test_fp64_check_in_the_mix(jobs, alone, gpu):
    import dataclasses

    checked = [dataclasses.replace(j, fp64_check="column") for j in jobs]
    print("continuous mix with float64 checks", flush=True)
    rep = SweepServer(max_active=4, chunk=64).run(checked, mode="continuous")
    for r in rep.records:
        assert _same(r.result, alone[r.job_id]), r.job_id
        assert r.fp64 is not None and r.fp64["mode"] == "column" and r.check_sets == 1
        assert len(r.fp64["columns"]) == 1 and 0.0 < r.fp64["max_abs_dS"] < 1e-4, r.fp64
        print(f"job {r.job_id}: fp32 vs fp64 {r.fp64['db']:.1f} dB")
    plain = SweepServer(max_active=4, chunk=64).run(jobs[:2]).records
    assert all(r.footprint < c.footprint for r, c in zip(plain, rep.records[:2]))


def test_packed_mix_equals_alone(jobs, alone, gpu):
    print("continuous mix, small jobs packed", flush=True)
    rep = SweepServer(max_active=4, chunk=64, pack_cells=40 ** 3).run(jobs, mode="continuous")
    assert sum(r.packed for r in rep.records) == len(jobs)
    for r in rep.records:
        assert _same(r.result, alone[r.job_id]), r.job_id
    print(f"packed: p50 {rep.p50:.3f} s  p99 {rep.p99:.3f} s  {rep.jobs_per_s:.2f} jobs/s")


def test_rejection_and_cancellation(jobs, alone, gpu):
    import threading

    first = SweepServer(max_active=8).run(jobs[:3]).records
    fps = sorted(r.footprint for r in first)
    budget = (fps[0] + fps[-1]) // 2  # the largest of the first three cannot fit
    big = max(first, key=lambda r: r.footprint).job_id
    fit = [j for j in jobs if j.job_id != big]
    early, resident = fit[-2].job_id, fit[-1].job_id
    srv = SweepServer(hbm_budget=budget, max_active=2, chunk=16)
    srv.cancel(early)  # before the run
    # the other is cancelled from another thread the moment it is admitted
    orig = srv._prepare

    def prepare(job, rec, t0, built=None):
        sw, solver, check = orig(job, rec, t0, built)
        if job.job_id == resident:
            setup = solver.setup

            def admitted():
                setup()
                t = threading.Thread(target=srv.cancel, args=(resident,))
                t.start()
                t.join()

            solver.setup = admitted
        return sw, solver, check

    srv._prepare = prepare
    print(f"rejection / cancellation, budget {budget / 1e6:.0f} MB", flush=True)
    rep = srv.run(jobs)
    by = {r.job_id: r for r in rep.records}
    assert by[big].status == "rejected" and by[big].result is None and "budget" in by[big].reason
    assert by[early].status == "cancelled" and by[early].result is None
    assert by[resident].status == "cancelled" and "resident" in by[resident].reason
    assert by[resident].result is None
    for r in rep.completed:
        assert _same(r.result, alone[r.job_id]), r.job_id
    assert len(rep.completed) == len(jobs) - 3 and all(r.status != "pending" for r in rep.records)


def test_budget_admission(jobs, alone, gpu):
    fps = sorted(r.footprint for r in SweepServer(max_active=8).run(jobs[:2]).records)
    budget = int(1.2 * max(fps))
    print(f"budget admission at {budget / 1e6:.0f} MB", flush=True)
    rep = SweepServer(hbm_budget=budget, max_active=8, chunk=64).run(jobs)
    assert rep.peak_bytes <= budget
    for r in rep.records:
        assert r.footprint <= budget
        assert _same(r.result, alone[r.job_id]), r.job_id


def test_sequential_mode(jobs, alone, gpu):
    rep = SweepServer(chunk=128).run(jobs, mode="sequential")
    assert rep.peak_resident == 1
    for r in rep.records:
        assert _same(r.result, alone[r.job_id]), r.job_id


def test_footprint_covers_cpml_and_monitors(jobs, gpu):
    import time

    import torch

    from magnelio_amd.boundaries import CPMLBoundary, PECBoundary
    from magnelio_amd.monitors import MonitorFieldFrequency
    from magnelio_amd.serve import JobRecord, SweepJob

    # a single guide with one port on z_min and a CPML on z_max
    base = next(j for j in jobs if j.build.tasks["mesh"][1].args[1] == 2)
    dims, _, objs = base.build.tasks["mesh"][1].args
    mesh, dt = bench._c5_mesh(dims, 2, objs)
    ports = [bench._c5_port(dims, 2, "port1", "zmin", 1, base.f_max)]
    bcs = {f: PECBoundary(f) for f in ("xmin", "xmax", "ymin", "ymax")}
    bcs["zmax"] = CPMLBoundary("zmax", mesh.grid, thickness_cells=8)
    job = SweepJob(99, 0.0, lambda: (mesh, ports, dt), base.n_steps, base.f_axis, base.f_max,
                   boundary_conditions=bcs,
                   monitors_factory=lambda: [MonitorFieldFrequency(np.linspace(8e9, 20e9, 41), fields=("E",))])
    srv = SweepServer()
    rec = JobRecord(99, 0.0)
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    sw, solver, check = srv._prepare(job, rec, time.perf_counter())
    assert check is None
    solver.setup()
    torch.cuda.synchronize()
    used = torch.cuda.memory_allocated() - before
    print(f"footprint {rec.footprint / 1e6:.1f} MB, allocated by setup {used / 1e6:.1f} MB", flush=True)
    assert rec.footprint >= used
    assert rec.footprint <= 1.1 * used + (64 << 20)
    mon = solver.monitors[0]
    assert mon.footprint(mesh.grid, bcs, solver.nsets) >= mon.bytes()
    # with the float64 check the footprint also covers the check solver
    import dataclasses

    rec2 = JobRecord(100, 0.0)
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated()
    _, solver2, check2 = srv._prepare(dataclasses.replace(job, job_id=100, fp64_check="column"),
                                      rec2, time.perf_counter())
    solver2.setup()
    check2.setup()
    torch.cuda.synchronize()
    used2 = torch.cuda.memory_allocated() - before
    print(f"checked: footprint {rec2.footprint / 1e6:.1f} MB, allocated {used2 / 1e6:.1f} MB", flush=True)
    assert rec2.footprint >= used2 and rec2.footprint > rec.footprint
