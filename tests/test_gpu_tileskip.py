"""Dead-unit skipping on the GPU equals dense launches (the reference's
dense-vs-skip identity, test_tile_skip_solver.py:70-86).

A guide with a thick metal wall (seeded aperture), metal bricks, a lossy
dielectric, modal ports on both z faces (K = 2 excitations) and a series-RLC
element whose edge chain crosses a metal brick (its units stay live):
fields and port records after 300 steps equal the dense run's, eagerly and
from graph replays, and a checkpoint resume (which zeroes the provably-zero
elements) continues identically."""

import numpy as np
import pytest
import torch

from magnelio_amd import mesh as mm
from magnelio_amd import ports as pp
from magnelio_amd.lumped import LumpedElement, SeriesRLC, build_lumped_element
from magnelio_amd.signals import Excitation
from magnelio_amd.solver import ChannelExcitation, FITTimeDomainSolver

pytestmark = pytest.mark.gpu

DIMS = (40, 18, 60)
SIZE = (22.86e-3, 10.16e-3, 34e-3)


def _case():
    axes = [np.linspace(0.0, L, n + 1) for L, n in zip(SIZE, DIMS)]
    g = mm.GridLines(*axes)
    W, H, L = SIZE
    rng = np.random.default_rng(8)
    wx, wy = rng.uniform(0.2, 0.5) * W, rng.uniform(0.2, 0.5) * H
    regions = [(mm.Material.pec(), (0.0, 0.0, 0.4 * L, W, H, 0.6 * L)),
               (mm.Material.air(), (wx, wy, 0.4 * L, wx + 0.3 * W, wy + 0.4 * H, 0.6 * L)),
               (mm.Material.pec(), (0.1 * W, 0.0, 0.2 * L, 0.3 * W, H, 0.3 * L)),
               (mm.Material.dielectric(5.0, sigma=0.05), (0.5 * W, 0.2 * H, 0.7 * L, 0.9 * W, 0.8 * H, 0.85 * L))]
    mesh = mm.Mesh.from_grid(g, regions=regions, pec_faces=("xmin", "xmax", "ymin", "ymax"))
    dt = mm.courant_dt(g, "normal")
    ports = [pp.build_port_local(n, f, mesh, dt=dt, f_calc=25e9, n_modes=2)
             for n, f in (("p1", "zmin"), ("p2", "zmax"))]
    x0 = axes[0][8]
    el = build_lumped_element(LumpedElement("r1", (x0, 0.0, axes[2][15]), (x0, H, axes[2][15]),
                                            SeriesRLC(R=30.0, L=1e-9)), mesh, mm.build_m_eps(mesh), dt)
    return mesh, ports, [el], dt


def _run(tile_skip, steps=300, use_graphs=False, split=None):
    mesh, ports, lumped, dt = _case()
    exc = [ChannelExcitation("p1", 0, Excitation(8e9, 20e9, 0)),
           ChannelExcitation("p2", 1, Excitation(15e9, 22e9, 1))]
    kw = dict(mesh=mesh, ports=ports, lumped=lumped, total_time_steps=steps, dt=dt, nsets=2,
              excitations=exc, tile_skip=tile_skip, use_graphs=use_graphs)
    sol = FITTimeDomainSolver(**kw)
    if split is None:
        sol.run()
        return sol
    sol.run(split)
    sd = sol.state_dict()
    sol2 = FITTimeDomainSolver(**kw)
    sol2.load_state_dict(sd)
    sol2.run()
    return sol2


def _same(a, b):
    assert torch.equal(a.e, b.e) and torch.equal(a.h, b.h)
    for s in range(2):
        ra, rb = a.recorded(s), b.recorded(s)
        assert ra.keys() == rb.keys()
        for k in ra:
            assert np.array_equal(ra[k][0], rb[k][0]) and np.array_equal(ra[k][1], rb[k][1])


def test_skip_equals_dense(gpu):
    print("tile skip: dense vs listed", flush=True)
    dense = _run("off")
    skip = _run("on")
    assert dense.tile_plan is None and skip.tile_plan is not None
    p = skip.tile_plan
    print(f"skipped units: E {p.skipped_e:.1%}  H {p.skipped_h:.1%}")
    assert p.skipped_e > 0.05 and p.skipped_h > 0.05
    assert float(dense.e.abs().max()) > 0
    _same(dense, skip)


def test_skip_graphs_and_resume(gpu):
    print("tile skip: graphs and resume", flush=True)
    dense = _run("off")
    _same(dense, _run("on", use_graphs=True))
    _same(dense, _run("on", split=140))
