"""Host-side set-up of the device ADE (no GPU): the dispersive blocks, their
edge sets, g and the folded W equal the oracle's; the pole table carries the
oracle's k, c in the field precision; a mesh without poles builds nothing."""

import numpy as np

from magnelio_amd import mesh as mm
from magnelio_amd.dispersion import DeviceDispersion
from oracle import dispersion as od
from oracle import mesh as om
from test_gpu_dispersion import ALL, BOXES, N, _axes, _boxes


def _mesh(names):
    axes = _axes()
    regions = [(mm.Material(name=b[0], epsilon=(b[1],) * 3, sigma=(b[2],) * 3, poles=b[3]), box)
               for b, box in _boxes(axes, names)]
    return axes, mm.Mesh.from_grid(mm.GridLines(*axes), regions=regions)


def test_blocks_match_oracle():
    axes, mesh = _mesh(ALL)
    x, y, z = axes
    dt = om.courant_dt(x, y, z, 0.95)
    dd = DeviceDispersion.from_mesh(mesh, dt, mm.pec_flat(mesh.pec_mask_edges, N))
    bx = _boxes(axes, ALL)
    mid = om.fill_aabb(x, y, z, [(m + 1, box) for m, (_, box) in enumerate(bx)])
    pec = om.bbox_walls(np.zeros(dd.W.size, bool), *N, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    ref = od.Dispersion(x, y, z, mid, {m + 1: list(b[3]) for m, (b, _) in enumerate(bx) if b[3]},
                        dt, pec_e=pec)
    assert len(dd.blocks) == len(ref.blocks) == 3
    for (_, idx, g, poles), rb in zip(dd.blocks, ref.blocks):
        assert np.array_equal(idx, rb.idx)
        assert np.array_equal(g, rb.g)
        for p, rp in zip(poles, rb.poles):
            assert p.weight == rp.weight and p.k == rp.k and p.c == rp.c
            row = p.row(np.float32)
            assert row[4] == np.float32(np.real(rp.k)) and row[6] == np.float32(np.real(rp.c))
            if p.complex:
                assert row[5] == np.float32(rp.k.imag) and row[7] == np.float32(rp.c.imag)
                assert row[2] == np.float32((1.0 + rp.k).real) and row[3] == np.float32((1.0 + rp.k).imag)
    assert np.array_equal(dd.W, ref.W)
    assert np.array_equal(dd.w_half, ref.w_half)


def test_no_poles_builds_nothing():
    names = ("lossy",)
    _, mesh = _mesh(names)
    assert DeviceDispersion.from_mesh(mesh, 1e-12, mm.pec_flat(mesh.pec_mask_edges, N)) is None
    assert all(not b[3] for b in BOXES if b[0] == "lossy")
