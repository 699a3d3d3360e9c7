"""Serving jobs (C5) in the continuous-batching mix against the oracle.

Three seeded sweep jobs from the C5 generator (bench.c5_jobs: a cross of two
WR-90 guides with 3 ports on its z and x ends x 3 modes, and two single
guides with one port x 2 modes; seeded dielectric bricks / spheres, every
channel excited), small grids (24^3-32^3 cells) and 300-400 steps, all
arriving at once so they run resident together on their own streams; the
oracle runs each job's excitations one at a time on cells filled from the
job's seeded geometry.  Every job's S-matrix (all channels x all
excitations, float32 engine) is within 1e-4 of the float64 oracle, and the
jobs' S-matrices do not depend on the mix (continuous == one job at a time,
bit for bit).
"""

import numpy as np
import pytest

import cases
from oracle import sparams as osp

pytestmark = pytest.mark.gpu


def _jobs():
    from bench import c5_jobs

    return c5_jobs(n_jobs=3, seed=17, rate=1e6, cells=(24 ** 3, 32 ** 3), steps=(300, 400), n_freq=21)


def test_c5_jobs_in_the_mix_vs_oracle(gpu):
    from magnelio_amd.serve import SweepServer

    jobs = _jobs()
    print("C5 jobs: " + ", ".join(
        f"{j.job_id}: dims {j.build.tasks['mesh'][1].args[0]} ports {j.build.tasks['mesh'][1].args[1]} "
        f"modes {j.build.tasks['port1'][0][3]} steps {j.n_steps}" for j in jobs), flush=True)
    rep = SweepServer(max_active=3).run(jobs, mode="continuous")
    assert rep.peak_resident >= 2
    seq = SweepServer(max_active=1).run(_jobs(), mode="sequential")
    for a, b in zip(rep.records, seq.records):
        assert np.array_equal(a.result.matrix, b.result.matrix, equal_nan=True)
    for job, rec in zip(jobs, rep.records):
        res = rec.result
        worst = 0.0
        for col, exc in enumerate(res.excitations):
            print(f"oracle job {job.job_id} excitation {exc}, {job.n_steps} steps (CPU)", flush=True)
            orc, ports, dt = cases.c5_job_oracle(job, exc)
            orc.run(job.n_steps)
            sig = {(p.name, m): (np.array(p.V[m]), np.array(p.I[m])) for p in ports
                   for m in range(p.n_modes)}
            lines = {(p.name, m): p.line[m] for p in ports for m in range(p.n_modes)}
            S = osp.s_params(sig, lines, tuple(exc), job.f_axis, dt)
            for row, ch in enumerate(res.channels):
                got = res.matrix[:, row, col]
                want = S[tuple(ch)]
                ok = np.isfinite(want) & np.isfinite(got)
                assert ok.sum() >= 0.8 * want.size, (job.job_id, exc, ch)
                worst = max(worst, float(np.max(np.abs(got[ok] - want[ok]))))
        print(f"  job {job.job_id}: max |S - S_oracle| {worst:.2e}", flush=True)
        assert worst <= 1e-4
