"""Write golden vectors of the thin-wire (m, 1/m) correction from the reference.

Run in the build container only (only the .npz this writes travels):

    python tests/golden/make_thinwire_fixtures.py /root/reference/src

The reference reaches its thin wires through CAD rasterisation (OCC,
absent here), but its OCC-free core ``apply_thin_wire_path``
(src/magnelio/mesh/_thin_wire.py:390-402, the form its own tests use) takes
a hand-built ``EdgePath``.  On a seeded graded 10 x 10 x 14 grid (adjacent
spacing ratio <= 1.25), two wires of radius 0.05 of the mean cell, applied
one after the other — a straight z wire at node (3, 4), k = 3..10, and an
L-shaped wire at node (6, 6): up along z from k = 2 to 9, then along +x to
i = 9 — whose PEC-masked edges the caller marks first.  Stored: the PEC
edge mask and the corrected M_eps / M_mu (build_M_eps / build_M_mu).
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
DIMS = (10, 10, 14)
WIRES = {
    "straight": [((3, 4, 3), (3, 4, 10))],
    "ell": [((6, 6, 2), (6, 6, 9)), ((6, 6, 9), (9, 6, 9))],
}


def axes():
    rng = np.random.default_rng(31)
    out = []
    for n in DIMS:
        w = np.exp(np.concatenate([[0.0], np.cumsum(rng.uniform(np.log(1 / 1.25), np.log(1.25), n - 1))]))
        w = w / w.mean() * 1e-3
        out.append(np.concatenate([[0.0], np.cumsum(w)]))
    return out


def radius():
    return 0.05 * 1e-3


def edges(legs):
    """[(axis, (i, j, k))] of the legs (lower-index base node per edge)."""
    out = []
    for a, b in legs:
        ax = next(t for t in range(3) if a[t] != b[t])
        lo, hi = sorted((a[ax], b[ax]))
        for s in range(lo, hi):
            ijk = list(a)
            ijk[ax] = s
            out.append((ax, tuple(ijk)))
    return out


def main(ref_src: str) -> None:
    sys.path.insert(0, ref_src)
    from magnelio._operators.material_matrices import build_M_eps, build_M_mu
    from magnelio.circuit.rasterize import EdgePath
    from magnelio.mesh._thin_wire import apply_thin_wire_path
    from magnelio.mesh.grid import GridLines
    from magnelio.mesh.indexing import edge_index_Ex, edge_index_Ey, edge_index_Ez
    from magnelio.mesh.mesher import Mesh

    x, y, z = axes()
    grid = GridLines(x=x, y=y, z=z)
    mesh = Mesh.from_grid(grid)
    nx, ny, nz = DIMS
    n_ex = nx * (ny + 1) * (nz + 1)
    n_ey = (nx + 1) * ny * (nz + 1)
    offs = (0, n_ex, n_ex + n_ey)
    fns = (edge_index_Ex, edge_index_Ey, edge_index_Ez)
    d = (grid.dx, grid.dy, grid.dz)
    for name, legs in WIRES.items():
        es = edges(legs)
        path = EdgePath(axes=["xyz"[ax] for ax, _ in es], ijk=[ijk for _, ijk in es],
                        signs=[1] * len(es), dls=[float(d[ax][ijk[ax]]) for ax, ijk in es],
                        flat_indices=[offs[ax] + fns[ax](*ijk, nx, ny, nz) for ax, ijk in es])
        for ax, ijk in es:
            mesh.pec_mask_edges[ax, fns[ax](*ijk, nx, ny, nz)] = True
        apply_thin_wire_path(mesh, path, radius(), name=name)
    np.savez_compressed(HERE / "thinwire_graded.npz", x=x, y=y, z=z, radius=np.float64(radius()),
                        pec=np.asarray(mesh.pec_mask_edges), m_eps=np.asarray(build_M_eps(mesh), float),
                        m_mu=np.asarray(build_M_mu(mesh), float))
    print("wrote thinwire_graded.npz")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
