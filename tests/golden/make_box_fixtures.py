"""Write golden vectors of a CPML box with a TF/SF plane wave and DFT monitors.

Run in the build container only (only the .npz this writes travels):

    python tests/golden/make_box_fixtures.py /root/reference/src

The reference's NumPy tier (``FITTimeDomainSolver``) on a reduced form of
the single-large-job case: a seeded graded 18 x 16 x 20 grid (adjacent
spacing ratio <= 1.3) with two dielectric bricks (eps_r 3 and 6), a 4-layer
``CPMLBoundary`` on all six faces (src/magnelio/boundaries/cpml.py,
defaults m 3, R 1e-8, kappa_max 7, alpha_max 0.02), a ``PlaneWaveSource``
(+z, x-polarised Gaussian, f_max 20 GHz) on [0.3L, 0.7L]^3
(sources/plane_wave.py:229-734), a ``MonitorFieldFrequency`` of E and H on a
region and a ``MonitorFarField`` Huygens box 2 cells inside the absorber
(monitors/field_frequency.py:249-294, far_field.py:123-248), 5 frequencies,
150 steps, dt = courant_dt(grid, "normal"), float64 and float32.  Stored:
e/h after 150 steps, every accumulator's bins, the CPML psi state; after
step 75 the solver's state_dict() and both monitors' result_dump()
(flattened "a/b" keys).
"""

from __future__ import annotations

import sys
import warnings
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from _flat import flatten  # noqa: E402
DIMS = (18, 16, 20)
SIZE = 24e-3
STEPS = 150
CHECKPOINT = 75  # the reference's state_dict() (psi, fields) is stored after this step
FREQS = np.array([4e9, 7e9, 10e9, 13e9, 16e9])
LAYERS = 4
FACES = ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax")


def axes():
    rng = np.random.default_rng(41)
    out = []
    for n in DIMS:
        w = np.exp(np.concatenate([[0.0], np.cumsum(rng.uniform(np.log(1 / 1.3), np.log(1.3), n - 1))]))
        w = w / w.sum() * SIZE
        out.append(np.concatenate([[0.0], np.cumsum(w)]))
    return out


BRICKS = [(3.0, (0.35 * SIZE, 0.35 * SIZE, 0.40 * SIZE, 0.55 * SIZE, 0.60 * SIZE, 0.55 * SIZE)),
          (6.0, (0.50 * SIZE, 0.40 * SIZE, 0.50 * SIZE, 0.65 * SIZE, 0.55 * SIZE, 0.62 * SIZE))]
TFSF = ((0.3 * SIZE,) * 3, (0.7 * SIZE,) * 3)
REGION = ((0.4 * SIZE, 0.3 * SIZE, 0.45 * SIZE), (0.6 * SIZE, 0.7 * SIZE, 0.6 * SIZE))


def main(ref_src: str) -> None:
    sys.path.insert(0, ref_src)
    from magnelio.boundaries.boundary_conditions import BoundaryConditions
    from magnelio.boundaries.cpml import CPMLBoundary
    from magnelio.materials.material import Material
    from magnelio.mesh.grid import GridLines
    from magnelio.mesh.mesher import Mesh
    from magnelio.monitors.far_field import MonitorFarField
    from magnelio.monitors.field_frequency import MonitorFieldFrequency
    from magnelio.solver.fit_td import FITTimeDomainSolver
    from magnelio.solver.stability import courant_dt
    from magnelio.sources.plane_wave import PlaneWaveSource

    x, y, z = axes()
    grid = GridLines(x=x, y=y, z=z)
    regions = [(Material(name=f"d{i}", epsilon=(e,) * 3), box) for i, (e, box) in enumerate(BRICKS)]
    out = {"x": x, "y": y, "z": z}
    for prec in ("double", "single"):
        mesh = Mesh.from_grid(grid, regions=regions,
                              boundary_conditions=BoundaryConditions(**{f: "CPML" for f in FACES}))
        mesh._pml_cells = {f: LAYERS for f in FACES}
        dt = courant_dt(grid, "normal")
        out["dt"] = np.float64(dt)
        bcs = {f: CPMLBoundary(f, grid, thickness_cells=LAYERS) for f in FACES}
        src = PlaneWaveSource(direction=(0.0, 0.0, 1.0), polarization=(1.0, 0.0, 0.0), corners=TFSF,
                              amplitude=1.0, waveform="gaussian", f_max=20e9)
        region = MonitorFieldFrequency(freqs=FREQS, corners=REGION, fields=["E", "H"], name="region")
        far = MonitorFarField(freqs=FREQS, name="far", margin_cells=2)
        sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, sources=[src],
                                  monitors=[region, far], total_time_steps=CHECKPOINT, dt=dt,
                                  verbose=False, backend="numpy", precision=prec,
                                  energy_stop_db=None, port_signal_stop_db=None)
        p = prec + "_"
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sol.run()
            out.update(flatten(sol.state_dict(), f"{p}state"))
            for mon in (region, far):
                out.update(flatten(mon.result_dump(), f"{p}dump{CHECKPOINT}_{mon.name}"))
            sol.total_time_steps = STEPS
            sol.run()
        out[p + "e"] = np.asarray(sol._fields.e_flat, np.float64).copy()
        out[p + "h"] = np.asarray(sol._fields.h_flat, np.float64).copy()
        for comp in region._components:
            out[f"{p}region_{comp}"] = np.asarray(region._accumulators[comp].result)
        r = region._region
        out["region_slices"] = np.array([[r.ix.start, r.ix.stop], [r.iy.start, r.iy.stop],
                                         [r.iz.start, r.iz.stop]])
        for bf in far._faces:
            for comp, acc in far._acc[bf.name].items():
                out[f"{p}far_{bf.name}_{comp}"] = np.asarray(acc.result)
        for face, bc in bcs.items():
            for k, v in bc.state_dict().items():
                out[f"{p}psi_{face}{k}"] = np.asarray(v, np.float64)
    np.savez_compressed(HERE / "box_tfsf_dft.npz", **out)
    print(f"wrote box_tfsf_dft.npz ({len(out)} arrays)")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
