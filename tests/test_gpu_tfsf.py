"""GPU parity of the TF/SF plane wave through the whole stepping path.

* The reference's own amplitude/leakage case (test_plane_wave.py:80-172):
  empty 24^3 PEC box, TF box inset 6 cells, all six propagation axes,
  FP64 — the HIP run equals the oracle bit for bit and meets the
  reference's gates (|inside - 1| < 1e-3, SF leakage < 1e-4).
* FP32 with CPML on all six faces, a dielectric brick straddling the TF
  box and a non-uniform grid: bit-identical to the oracle after 150 steps,
  eager and graph-replayed.
"""

import numpy as np
import pytest

from magnelio_amd import mesh as mm
from magnelio_amd.boundaries import CPMLBoundary, PECBoundary
from magnelio_amd.solver import FITTimeDomainSolver
from magnelio_amd.sources import PlaneWaveSource
from oracle import cpml as oc
from oracle import fdtd as ofd
from oracle import mesh as om
from oracle import tfsf as ot
from oracle import yee
from test_oracle_tfsf import CASES, tfsf_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("direction,polarization", CASES)
def test_reference_amplitude_case_bit_exact(direction, polarization, gpu):
    n, inset = 24, 6
    print(f"TF/SF 24^3 fp64 direction={direction}", flush=True)
    orc, n_steps, axes = tfsf_case(direction, polarization, n, inset)
    orc.run(n_steps)
    grid = mm.GridLines(*axes)
    mesh = mm.Mesh.from_grid(grid)
    src = PlaneWaveSource(direction=direction, polarization=polarization,
                          corners=(tuple(a[inset] for a in axes), tuple(a[n - inset] for a in axes)),
                          f_max=6e9, waveform="gaussian")
    bcs = {f: PECBoundary(f) for f in mm.FACES}
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, sources=[src],
                              total_time_steps=n_steps, dt=mm.courant_dt(grid, "normal"),
                              precision="double")
    assert sol.dt == orc.dt
    sol.run()
    e, _ = sol.fields(0)
    assert np.array_equal(e, orc.e)
    comp = int(np.argmax(np.abs(polarization)))
    arr = np.abs(yee.split(e, n, n, n, "E")[comp]) / (24e-3 / n)
    core = slice(inset + 1, n - inset)
    inside = float(arr[core, core, core].max())
    shell = arr.copy()
    shell[inset:-inset, inset:-inset, inset:-inset] = 0.0
    outside = float(shell[1:-1, 1:-1, 1:-1].max())
    print(f"  inside {inside:.6f} leakage {outside:.2e}", flush=True)
    assert inside == pytest.approx(1.0, abs=1e-3)
    assert outside < 1e-4


def _cpml_case(dtype):
    rng = np.random.default_rng(21)
    dims = (22, 18, 26)
    x, y, z = (np.cumsum(np.r_[0.0, rng.uniform(0.8e-3, 1.2e-3, n)]) for n in dims)
    brick = (x[9], y[7], z[11], x[15], y[13], z[20])
    mid = om.fill_aabb(x, y, z, [(1, brick)])
    eps = [[1.0, 4.0]] * 3
    me = om.mass_e(x, y, z, mid, eps, om.EPS0)
    mmu = om.mass_h(x, y, z, mid, [[1.0, 1.0]] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, np.zeros(me.size, bool))
    ah, bh = om.alpha_beta(mmu, np.zeros_like(mmu), dt)
    faces = ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"]
    nl = 5
    ofaces = [oc.CPMLFace(f, x, y, z, dt, dtype, n_layers=nl) for f in faces]
    corners = ((x[7], y[6], z[8]), (x[16], y[13], z[19]))
    box = ot.snap_box(x, y, z, corners)
    src = ot.PlaneWave(x, y, z, be, bh, dt, direction=(0, 0, 1), polarization=(1, 0, 0), box=box,
                       f_max=30e9, dtype=dtype)
    orc = ofd.Oracle(dims, ae, be, ah, bh, dt, cpml=ofaces, sources=[src], dtype=dtype)
    return (x, y, z), brick, faces, nl, corners, orc


@pytest.mark.parametrize("use_graphs", [False, True])
def test_cpml_dielectric_fp32_bit_exact(use_graphs, gpu):
    print(f"TF/SF + CPML fp32 graphs={use_graphs}", flush=True)
    axes, brick, faces, nl, corners, orc = _cpml_case(np.float32)
    steps = 150
    orc.run(steps)
    grid = mm.GridLines(*axes)
    diel = mm.Material.dielectric(4.0)
    mesh = mm.Mesh.from_grid(grid, regions=[(diel, brick)], pec_faces=())
    bcs = {f: CPMLBoundary(f, grid, thickness_cells=nl) for f in faces}
    src = PlaneWaveSource(direction=(0, 0, 1), polarization=(1, 0, 0), corners=corners, f_max=30e9)
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, sources=[src],
                              total_time_steps=steps, dt=mm.courant_dt(grid, "normal"),
                              precision="single", use_graphs=use_graphs)
    sol.run()
    if use_graphs:
        assert sol._graph is not None, getattr(sol, "graph_error", "graph not captured")
    e, h = sol.fields(0)
    assert np.abs(orc.e).max() > 0
    assert np.array_equal(e, orc.e)
    assert np.array_equal(h, orc.h)
