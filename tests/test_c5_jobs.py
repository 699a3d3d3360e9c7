"""The serving workload's generator and its oracle geometry (CPU).

* the seeded stream holds 1-4 port jobs: a WR-90 guide with z-end ports for
  1-2 ports, a cross of two WR-90 guides (PEC corner blocks) with z- and
  x-end ports for 3-4;
* a cross job's product mesh (materials, PEC mask with the walls and the
  port-plane flattening) equals the oracle's restatement of the job bit for
  bit;
* its product ports (numerical TE modes on each face's window, built on the
  feed slab) equal the oracle's closed-form TE modes on the same windows
  (cut-offs 1e-10, profiles 1e-9 of their scale).
"""

import numpy as np
import pytest

import bench
import cases
from magnelio_amd import mesh as mm


def _jobs():
    return bench.c5_jobs(n_jobs=12, seed=17, rate=1e6, cells=(24 ** 3, 40 ** 3), steps=(300, 400), n_freq=21)


def test_stream_covers_one_to_four_ports():
    jobs = bench.c5_jobs(n_jobs=200)
    ports = [j.build.tasks["mesh"][1].args[1] for j in jobs]
    assert set(ports) == {1, 2, 3, 4}
    for j in jobs:
        dims, n_ports, _ = j.build.tasks["mesh"][1].args
        faces = sorted(v[0][2] for k, v in j.build.tasks.items() if k.startswith("port"))
        assert faces == sorted(bench.C5_FACES[:n_ports])
        if n_ports > 2:
            assert dims[0] == dims[2] and dims[0] % 16 == 0


def _cross_job():
    return next(j for j in _jobs() if j.build.tasks["mesh"][1].args[1] == 4)


def test_cross_job_mesh_equals_oracle():
    job = _cross_job()
    dims, n_ports, objs = job.build.tasks["mesh"][1].args
    mesh, dt = bench._c5_mesh(dims, n_ports, objs)
    me = mm.build_m_eps(mesh)
    pec = mesh.pec_mask_edges
    for f in bench.C5_FACES[:n_ports]:
        me = mm.flatten_port_mass(me, dims, f)
        pec = mm.flatten_port_pec(pec, dims, f)
    orc, ports, dt_o = cases.c5_job_oracle(job, ("port1", 0))
    assert dt == dt_o
    co = mm.update_coefficients(me, mm.build_m_mu(mesh), mm.build_m_sigma(mesh), mm.build_m_sigma_m(mesh), dt,
                                mm.pec_flat(pec, dims))
    assert np.array_equal(co.alpha_e, orc.ae) and np.array_equal(co.beta_e, orc.be)


def test_cross_job_ports_equal_closed_form():
    job = _cross_job()
    mesh, ports, dt = job.build()
    _, oports, _ = cases.c5_job_oracle(job, ("port1", 0))
    assert [p.face for p in ports] == list(bench.C5_FACES[:4])
    for p, o in zip(ports, oports):
        assert p.termination_kinds == ["dtbc"] * p.n_modes
        for pm, om_ in zip(p.modes, o.modes):
            assert pm.omega_c == pytest.approx(om_["wc"], rel=1e-10)
            sc = max(np.abs(om_["e_u"]).max(), np.abs(om_["e_v"]).max())
            np.testing.assert_allclose(pm.e_u, om_["e_u"], rtol=0, atol=1e-9 * sc)
            np.testing.assert_allclose(pm.e_v, om_["e_v"], rtol=0, atol=1e-9 * sc)
        np.testing.assert_allclose(p.record_scale, o.record_scale, rtol=1e-9)
