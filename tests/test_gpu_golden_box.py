"""The HIP engine on the reference's CPML / TF-SF / DFT box run
(tests/golden/box_tfsf_dft.npz, tests/golden/make_box_fixtures.py): a
reduced form of the single-large-job case — graded grid, dielectric bricks,
4-layer CPML on six faces inside the curl launches, +z TF/SF Gaussian, a
region E/H DFT monitor and a far-field Huygens box — 150 steps.

* float64: fields within 1e-10 of the reference's; float32: within 1e-5
  of the reference's float32 run;
* region and far-field bins (pipelined and whole-period contractions)
  within 1e-6 of the peak of the reference's bins of the same precision:
  the phasor contraction runs on float32 MFMA over float32 samples in
  either precision (products rounded at ~6e-8; float64 or compensated bins
  accumulate them), so the bins carry ~1e-7 of the peak; the float32 run's
  bins also within 1e-5 of the reference's float64 bins.
"""

import numpy as np
import pytest

import cases
from magnelio_amd import mesh as mm
from magnelio_amd.boundaries import CPMLBoundary
from magnelio_amd.monitors import TANGENTIALS, MonitorFarField, MonitorFieldFrequency
from magnelio_amd.solver import FITTimeDomainSolver
from magnelio_amd.sources import PlaneWaveSource

pytestmark = pytest.mark.gpu


def _run(precision, pipelined):
    f, mk = cases.golden_box()
    g = mm.GridLines(f["x"], f["y"], f["z"])
    mesh = mm.Mesh.from_grid(g, regions=[(mm.Material.dielectric(e), box) for e, box in mk.BRICKS],
                             pec_faces=())
    bcs = {fc: CPMLBoundary(fc, g, thickness_cells=mk.LAYERS) for fc in mk.FACES}
    src = PlaneWaveSource(direction=(0, 0, 1), polarization=(1, 0, 0), corners=mk.TFSF, f_max=20e9)
    region = MonitorFieldFrequency(mk.FREQS, corners=mk.REGION, fields=("E", "H"), block=64,
                                   pipelined=pipelined)
    far = MonitorFarField(mk.FREQS, margin_cells=2, block=64, pipelined=pipelined)
    dt = mm.courant_dt(g, "normal")
    assert dt == float(f["dt"])
    sol = FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, sources=[src], monitors=[region, far],
                              total_time_steps=mk.STEPS, dt=dt, precision=precision, record=False)
    sol.run()
    return f, mk, sol, region, far


def _far_bins(far):
    """{(face, comp): (F, n1, n2)} from the product monitor's groups."""
    res = far.result(0)
    seen = {}
    out = {}
    for face, axis, _, _, _ in far.faces:
        for comp in TANGENTIALS[axis]:
            k = seen.get(comp, 0)
            seen[comp] = k + 1
            v = res[comp]
            out[(face, comp)] = v[k] if isinstance(v, list) else v
    return out


@pytest.mark.parametrize("precision,pipelined", [("double", True), ("double", False),
                                                 ("single", True)])
def test_hip_box_matches_reference(precision, pipelined, gpu):
    print(f"HIP golden box {precision} pipelined={pipelined}", flush=True)
    f, mk, sol, region, far = _run(precision, pipelined)
    p = precision + "_"
    e, h = sol.fields(0)
    gate = 1e-10 if precision == "double" else 1e-5
    for comp, got in (("e", e), ("h", h)):
        want = f[p + comp]
        err = float(np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want))
        print(f"  {comp}: rel {err:.2e}", flush=True)
        assert err <= gate, comp
    bgate = 1e-6
    rs = f["region_slices"]
    assert tuple((s.start, s.stop) for s in region.slices) == tuple(tuple(int(v) for v in r) for r in rs)
    got = region.result(0)
    bins = _far_bins(far)
    assert len(bins) == 24
    for ref, g in ((p, bgate), ("double_", 1e-5)):
        worst = 0.0
        for c, b in got.items():
            want = f[f"{ref}region_{c}"]
            assert b.shape == want.shape, c
            worst = max(worst, float(np.abs(b - want).max() / np.abs(want).max()))
        print(f"  region bins vs the reference's {ref[:-1]}: max err / peak {worst:.2e}", flush=True)
        assert worst <= g
        peak = max(np.abs(f[k]).max() for k in f if k.startswith(ref + "far_"))
        worst = 0.0
        for (face, c), b in bins.items():
            want = f[f"{ref}far_{face}_{c}"]
            assert b.shape == want.shape, (face, c)
            worst = max(worst, float(np.abs(b - want).max() / peak))
        print(f"  far-field bins vs the reference's {ref[:-1]}: max err / peak {worst:.2e}", flush=True)
        assert worst <= g
