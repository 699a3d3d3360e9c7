"""Thin-wire correction against the reference's own values (golden vectors).

tests/golden/thinwire_graded.npz (tests/golden/make_thinwire_fixtures.py):
the reference's OCC-free ``apply_thin_wire_path``
(src/magnelio/mesh/_thin_wire.py:390-402) on a seeded graded grid with a
straight and an L-shaped wire (0.05 of the mean cell), then build_M_eps /
build_M_mu.  The product's ``apply_thin_wires`` with the same node
polylines gives the same PEC edge mask and the same corrected masses bit for
bit: the ring stencil, r0 from the geometric mean, the per-face / per-edge
minimum collapse at the corner and the category-1/2 write order.
"""

import sys
import warnings
from pathlib import Path

import numpy as np

from magnelio_amd import mesh as mm
from magnelio_amd.thinwire import ThinWire, apply_thin_wires

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_thinwire_fixtures as mk  # noqa: E402


def test_thin_wire_masses_equal_reference():
    f = np.load(GOLDEN / "thinwire_graded.npz")
    g = mm.GridLines(f["x"], f["y"], f["z"])
    mesh = mm.Mesh.from_grid(g, pec_faces=mm.FACES)  # the reference closes every face
    wires = []
    for name, legs in mk.WIRES.items():
        pts = [legs[0][0]] + [b for _, b in legs]
        wires.append(ThinWire(name, tuple((g.x[i], g.y[j], g.z[k]) for i, j, k in pts),
                              float(f["radius"])))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # anisotropy warnings, as the reference's
        apply_thin_wires(mesh, wires)
    assert np.array_equal(np.asarray(mesh.pec_mask_edges), f["pec"])
    me, mu = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    assert np.array_equal(me, f["m_eps"])
    assert np.array_equal(mu, f["m_mu"])
    # the correction did act: some faces and edges differ from the bare grid
    bare = mm.Mesh.from_grid(g, pec_faces=mm.FACES)
    assert np.sum(mm.build_m_mu(bare) != mu) >= 8
    assert np.sum(mm.build_m_eps(bare) != me) >= 8
