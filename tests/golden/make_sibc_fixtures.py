"""Write golden vectors of surface-impedance (SIBC) walls from the reference.

Run in the build container only (only the .npz this writes travels):

    python tests/golden/make_sibc_fixtures.py /root/reference/src

A seeded graded 10 x 8 x 12 cavity (adjacent spacing ratio <= 1.3) closed
by PEC on all six faces, holding a lossy-metal brick (``Material.
lossy_metal``, sigma 5.8e5 S/m) and a plain PEC brick; the walls are
enumerated by the reference's ``enumerate_sibc_surfaces`` (staircase
solids + the six boundary faces, src/magnelio/mesh/_surfaces.py:935-1030),
resolved with a 2e6 S/m fallback for the walls without their own
conductor (``resolve_wall_conductors``) and fitted over 1-100 GHz
(``fit_wall_impedances``, materials/surface_impedance.py:333-383).  The
reference's NumPy tier then marches a seeded random E field (PEC edges
zero) with ``sibc=SIBCSpec(...)`` for 300 steps at courant_dt(grid,
"normal"), float64 and float32.  Stored: every surface (tag, comp,
flat_idx, weight, inv_l_dual), every fit (c0, branches), the solver's
alpha_H / beta_H, e/h after steps 1, 100 and 300 and its state_dict()
after step 100 (flattened "a/b" keys).
"""

from __future__ import annotations

import sys
import warnings
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from _flat import flatten  # noqa: E402

DIMS = (10, 8, 12)
SIZE = 10e-3
SNAPS = (1, 100, 300)
CHECKPOINT = 100
FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")
SIGMA_CU = 5.8e5
SIGMA_FALLBACK = 2.0e6
F_LO, F_HI = 1e9, 1e11


def axes():
    rng = np.random.default_rng(43)
    out = []
    for n in DIMS:
        w = np.exp(np.concatenate([[0.0], np.cumsum(rng.uniform(np.log(1 / 1.3), np.log(1.3), n - 1))]))
        w = w / w.sum() * SIZE
        out.append(np.concatenate([[0.0], np.cumsum(w)]))
    return out


def boxes(x, y, z):
    """(lossy brick, PEC brick) corners on grid lines."""
    return ((x[2], y[2], z[3], x[5], y[5], z[6]), (x[6], y[3], z[7], x[8], y[6], z[9]))


def e0(n_e, pec_flat, seed=11):
    e = np.random.default_rng(seed).standard_normal(n_e)
    e[pec_flat] = 0.0
    return e


def main(ref_src: str) -> None:
    sys.path.insert(0, ref_src)
    from magnelio.boundaries.pec import PECBoundary
    from magnelio.materials.material import Material
    from magnelio.materials.surface_impedance import fit_wall_impedances
    from magnelio.mesh._surfaces import enumerate_sibc_surfaces, resolve_wall_conductors
    from magnelio.mesh.grid import GridLines
    from magnelio.mesh.mesher import Mesh
    from magnelio.solver._sibc import SIBCSpec
    from magnelio.solver.fit_td import FITTimeDomainSolver
    from magnelio.solver.stability import courant_dt

    x, y, z = axes()
    grid = GridLines(x=x, y=y, z=z)
    lossy, plain = boxes(x, y, z)
    regions = [(Material.lossy_metal("cu", sigma=SIGMA_CU), lossy), (Material.pec(), plain)]
    mesh = Mesh.from_grid(grid, regions=regions)
    surfs = enumerate_sibc_surfaces(mesh, bc_pec_faces=FACES)
    resolved = resolve_wall_conductors(mesh, surfs, sigma=SIGMA_FALLBACK)
    fits = fit_wall_impedances(resolved, F_LO, F_HI)
    spec = SIBCSpec(surfaces=tuple(surfs), fits=fits)
    out = {"x": x, "y": y, "z": z, "material_id": np.asarray(mesh.material_id, np.int32),
           "pec_cell_ids": np.array(sorted(m for m, mat in mesh.material_library.items() if mat.is_pec)),
           "n_surfaces": np.int64(len(surfs))}
    pec = mesh.pec_mask_edges
    nx, ny, nz = DIMS
    n_ex, n_ey, n_ez = nx * (ny + 1) * (nz + 1), (nx + 1) * ny * (nz + 1), (nx + 1) * (ny + 1) * nz
    pec_flat = np.concatenate([pec[0, :n_ex], pec[1, :n_ey], pec[2, :n_ez]]).astype(bool)
    out["pec_edges"] = pec_flat
    for i, s in enumerate(surfs):
        out[f"surf{i}_tag"] = np.array(str(s.tag))
        out[f"surf{i}_comp"] = np.asarray(s.comp)
        out[f"surf{i}_flat_idx"] = np.asarray(s.flat_idx)
        out[f"surf{i}_weight"] = np.asarray(s.weight)
        out[f"surf{i}_inv_l_dual"] = np.asarray(s.inv_l_dual)
        fit = fits[s.tag]
        out[f"surf{i}_c0"] = np.float64(fit.c0)
        out[f"surf{i}_branches"] = np.array(fit.branches, np.float64).reshape(-1, 2)
    dt = courant_dt(grid, "normal")
    out["dt"] = np.float64(dt)
    for prec in ("double", "single"):
        sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions={f: PECBoundary(f) for f in FACES},
                                  total_time_steps=SNAPS[0], dt=dt, verbose=False, backend="numpy",
                                  precision=prec, energy_stop_db=None, port_signal_stop_db=None,
                                  sibc=spec)
        p = prec + "_"
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sol.setup()
            sol._fields.e_flat[:] = e0(n_ex + n_ey + n_ez, pec_flat)
            out[p + "alpha_h"] = np.asarray(sol._alpha_H, np.float64)
            out[p + "beta_h"] = np.asarray(sol._beta_H, np.float64)
            for n in SNAPS:
                sol.total_time_steps = n
                sol.run()
                out[f"{p}e{n}"] = np.asarray(sol._fields.e_flat, np.float64).copy()
                out[f"{p}h{n}"] = np.asarray(sol._fields.h_flat, np.float64).copy()
                if n == CHECKPOINT:
                    out.update(flatten(sol.state_dict(), f"{p}state"))
    np.savez_compressed(HERE / "sibc_cavity.npz", **out)
    print(f"wrote sibc_cavity.npz ({len(out)} arrays, {len(surfs)} surfaces)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
