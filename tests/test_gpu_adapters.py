"""A solver assembled from the reference's objects runs on the GPU.

The WR-90 guide of tests/golden/ports_wr90.npz: mesh, PEC boundaries and
both modal ports as reference-shaped stand-ins (the ports carry the values
the reference's PortOperatorModal held), port 1 mode 0 excited the way the
reference's operator holds it (waveform times source_scale), through
``adapters.solver_from_reference``; 300 float64 steps equal the run of this
backend's own builders within 1e-10, V/I within 1e-9.
"""

import numpy as np
import pytest

from test_adapters import _operator_standin

pytestmark = pytest.mark.gpu


def test_adapter_built_solver_runs(gpu):
    from pathlib import Path
    from types import SimpleNamespace as NS

    from magnelio_amd import adapters
    from magnelio_amd import mesh as mm
    from magnelio_amd import ports as pp
    from magnelio_amd.boundaries import PECBoundary
    from magnelio_amd.signals import resolve_excitation
    from magnelio_amd.solver import ChannelExcitation, FITTimeDomainSolver

    f = dict(np.load(Path(__file__).resolve().parent / "golden" / "ports_wr90.npz"))
    g = mm.GridLines(f["x"], f["y"], f["z"])
    walls = ("xmin", "xmax", "ymin", "ymax")
    own_mesh = mm.Mesh.from_grid(g, pec_faces=walls)
    dt = mm.courant_dt(g, "normal")
    me, mu = mm.build_m_eps(own_mesh), mm.build_m_mu(own_mesh)
    own_ports = [pp.build_port(n, fc, own_mesh, me, mu, dt=dt, f_calc=25e9, n_modes=3)
                 for n, fc in (("port1", "zmin"), ("port2", "zmax"))]
    wf = resolve_excitation(own_ports[0].modes[0].omega_c, 0.0, 25e9, 0)
    own = FITTimeDomainSolver(mesh=own_mesh, boundary_conditions={w: PECBoundary(w) for w in walls},
                              ports=own_ports, total_time_steps=300, dt=dt, precision="double",
                              excitations=[ChannelExcitation("port1", 0, wf)])
    own.run()

    lib = {k: NS(name=m.name, epsilon=m.epsilon, mu=m.mu, sigma=m.sigma, sigma_m=m.sigma_m,
                 is_pec=m.is_pec, dispersion=None, dispersion_mu=None)
           for k, m in own_mesh.material_library.items()}
    ref_mesh = NS(grid=NS(x=g.x, y=g.y, z=g.z), material_id=own_mesh.material_id, material_library=lib,
                  pec_mask_edges=own_mesh.pec_mask_edges.copy(),
                  boundary_conditions={w: "PEC" for w in walls} | {"zmin": "PMC", "zmax": "PMC"})
    ops = [_operator_standin(f, "port1"), _operator_standin(f, "port2")]
    scale = float(ops[0]._source_scale[0])
    ops[0]._excitation_mode = 0
    ops[0]._excitation_waveform = lambda t: scale * wf(t)
    ops[1]._excitation_mode = None
    PECRef = type("PECBoundary", (), {})
    sol = adapters.solver_from_reference(mesh=ref_mesh, boundary_conditions={w: PECRef() for w in walls},
                                         ports=ops, total_time_steps=300, dt=dt, precision="double",
                                         f_calc=25e9)
    sol.run()
    e, h = sol.fields(0)
    e0, h0 = own.fields(0)
    err = max(np.linalg.norm(e - e0) / np.linalg.norm(e0), np.linalg.norm(h - h0) / np.linalg.norm(h0))
    print(f"adapter-built vs own-built solver, 300 float64 steps: rel {err:.2e}", flush=True)
    assert err <= 1e-10
    a, b = sol.recorded(0), own.recorded(0)
    vmax = max(np.abs(v[0]).max() for v in b.values())
    for k in b:
        assert np.max(np.abs(a[k][0] - b[k][0])) <= 1e-9 * vmax
