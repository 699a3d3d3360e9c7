"""The feed-slab port build equals the full-grid build bit for bit (CPU):
C1 guide ports on both z faces with 3 modes, and a loaded guide whose
dielectric sits outside the feed, ports on x faces with a single wall face
left out."""

import dataclasses

import numpy as np

import cases
from magnelio_amd import mesh as mm
from magnelio_amd import ports as pp


def _same_model(a, b):
    for f in dataclasses.fields(a):
        va, vb = getattr(a, f.name), getattr(b, f.name)
        if f.name == "plane":
            for g in dataclasses.fields(va):
                x, y = getattr(va, g.name), getattr(vb, g.name)
                assert np.array_equal(np.asarray(x), np.asarray(y)), g.name
        elif f.name == "modes":
            for ma, mb in zip(va, vb):
                for g in dataclasses.fields(ma):
                    assert np.array_equal(np.asarray(getattr(ma, g.name)), np.asarray(getattr(mb, g.name))), g.name
        elif isinstance(va, (list, tuple)):
            assert len(va) == len(vb)
            for x, y in zip(va, vb):
                assert (x is None and y is None) or np.array_equal(np.asarray(x), np.asarray(y)), f.name
        else:
            assert np.array_equal(np.asarray(va), np.asarray(vb)), f.name


def test_c1_ports_local_equal_full():
    mesh, ports, dt = cases.c1_mesh_ports()
    for P in ports:
        L = pp.build_port_local(P.name, P.face, mesh, dt=dt, f_calc=cases.C1_FMAX, n_modes=3)
        _same_model(L, P)


def test_loaded_x_ports_local_equal_full():
    x = np.linspace(0, 30e-3, 31)
    y = np.linspace(0, 10e-3, 11)
    z = np.concatenate([[0.0], np.cumsum(np.linspace(0.9e-3, 1.1e-3, 14))])
    g = mm.GridLines(x, y, z)
    diel = mm.Material.dielectric(6.0, sigma=0.2)
    mesh = mm.Mesh.from_grid(g, regions=[(diel, (12e-3, 2e-3, 3e-3, 18e-3, 8e-3, 9e-3))],
                             pec_faces=("ymin", "ymax", "zmin", "zmax"))
    dt = mm.courant_dt(g, "normal")
    me, mu = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    for face in ("xmin", "xmax"):
        full = pp.build_port("p", face, mesh, me, mu, dt=dt, f_calc=40e9, n_modes=2)
        local = pp.build_port_local("p", face, mesh, dt=dt, f_calc=40e9, n_modes=2)
        _same_model(local, full)
