"""Thin-wire corrections (CPU): closed forms of the (m, 1/m) encoding on a
uniform grid, the conservative composition at a corner, PEC masking of the
chain, and the radius validation (_thin_wire.py:76-330)."""

import math

import numpy as np
import pytest

from magnelio_amd import mesh as mm
from magnelio_amd.solver import FITTimeDomainSolver
from magnelio_amd.thinwire import KAPPA0, ThinWire, apply_thin_wires, rasterize

D = 1e-3


def _mesh(n=(12, 12, 20)):
    g = mm.GridLines(*(np.linspace(0.0, D * k, k + 1) for k in n))
    return mm.Mesh.from_grid(g)


def test_straight_wire_closed_form():
    mesh = _mesh()
    bare_e, bare_h = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    a = 0.05 * D
    wire = ThinWire("w", ((6 * D, 6 * D, 2 * D), (6 * D, 6 * D, 15 * D)), a)
    apply_thin_wires(mesh, [wire])
    wc = mesh.wire_corrections
    m = math.log(D / a) / math.log(1.0 / KAPPA0)
    assert wc.face_idx.size == 4 * 13 and np.allclose(wc.face_m, m, rtol=1e-14)
    e, h = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    np.testing.assert_allclose(h[wc.face_idx], m * bare_h[wc.face_idx], rtol=1e-14)
    np.testing.assert_allclose(e[wc.edge_idx], bare_e[wc.edge_idx] / wc.edge_m, rtol=1e-14)
    untouched_h = np.setdiff1d(np.arange(h.size), wc.face_idx)
    untouched_e = np.setdiff1d(np.arange(e.size), wc.edge_idx)
    assert np.array_equal(h[untouched_h], bare_h[untouched_h])
    assert np.array_equal(e[untouched_e], bare_e[untouched_e])
    # the chain is PEC and the mask is no longer the plain OR rule
    nx, ny, nz = mesh.dims
    pec = mm.pec_flat(mesh.pec_mask_edges, mesh.dims)
    n_ex, n_ey = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1)
    chain = [n_ex + n_ey + (6 * (ny + 1) + 6) * nz + k for k in range(2, 15)]
    assert pec[chain].all() and not mesh.pec_rule
    assert not FITTimeDomainSolver(mesh=mesh)._device_coefficients()
    # 2 radial edges per face, shared between consecutive segments
    assert wc.edge_idx.size == 4 * 14


def test_corner_keeps_smaller_factor():
    mesh = _mesh()
    wire = ThinWire("L", ((4 * D, 6 * D, 3 * D), (4 * D, 6 * D, 10 * D), (9 * D, 6 * D, 10 * D)), 0.05 * D)
    path = rasterize(wire, mesh.grid)
    assert len(path) == 7 + 5 and path[0][0] == 2 and path[-1][0] == 0
    apply_thin_wires(mesh, [wire])
    assert np.isfinite(mesh.wire_corrections.face_m).all()


def test_radius_validation():
    mesh = _mesh()
    with pytest.raises(ValueError):
        apply_thin_wires(mesh, [ThinWire("fat", ((6 * D, 6 * D, 2 * D), (6 * D, 6 * D, 8 * D)), 0.31 * D)])
    with pytest.raises(ValueError):
        rasterize(ThinWire("off", ((6.5 * D, 6 * D, 2 * D), (6.5 * D, 6 * D, 8 * D)), 0.05 * D), mesh.grid)
    with pytest.warns(UserWarning):
        apply_thin_wires(_mesh(), [ThinWire("warm", ((6 * D, 6 * D, 2 * D), (6 * D, 6 * D, 8 * D)), 0.25 * D)])
