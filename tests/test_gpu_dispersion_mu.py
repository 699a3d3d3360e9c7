"""GPU parity of the ADE on the H side: mu(w) pole-residue materials
(src/magnelio/solver/_dispersion.py:45-76, fit_td.py:430-461, :826-836,
:891-921).

A graded PEC box holding mu-dispersive blocks (Debye, Drude with a magnetic
loss, a Lorentz conjugate pair) next to an eps-dispersive one, started from
a seeded random E field, against the oracle's restatement (the same
operator over the H faces, g = MU0 A_primal / L_dual):

* real poles only, FP64 and FP32: fields bit-identical;
* with the conjugate pairs, FP64 and FP32: bit-identical when the host's
  NumPy rounds complex products with FMA (the rounding the kernel follows),
  else within 1e-12 / 1e-5 relative L2;
* with a CPML on z_max (per-face passes after the mu(w) completion): FP32
  bit-identical;
* K = 2 sets equal their single-set runs; a state_dict round trip mid-run
  (its "dispersion_mu" entry) resumes to the same fields.
"""

import math

import numpy as np
import pytest
import torch

from magnelio_amd import layout
from magnelio_amd import mesh as mm
from magnelio_amd.boundaries import CPMLBoundary, PECBoundary
from magnelio_amd.solver import FITTimeDomainSolver
from oracle import cpml as oc
from oracle import dispersion as od
from oracle import fdtd as ofd
from oracle import mesh as om
from test_gpu_dispersion import _numpy_complex_fma

pytestmark = pytest.mark.gpu

N = (14, 12, 12)
TAU = 1.0 / (2 * math.pi * 5e9)
WPM, GM = 2 * math.pi * 10e9, 2 * math.pi * 4e9
LOR_MU = od.lorentz_pair(0.9, 2 * math.pi * 14e9, 2 * math.pi * 2e9)
LOR_EPS = od.lorentz_pair(1.4, 2 * math.pi * 17e9, 2 * math.pi * 3e9)
# (name, eps_inf, sigma, eps poles, mu_inf, sigma_m, mu poles, box in cells)
BOXES = [
    ("mu_debye", 1.0, 0.0, (), 1.8, 0.0, ((-1.0 / TAU, 1.2 / TAU),), (2, 2, 2, 7, 6, 6)),
    ("mu_drude", 2.0, 0.01, (), 1.0, 150.0, ((0.0, WPM * WPM / GM), (-GM, -WPM * WPM / GM)),
     (7, 6, 5, 12, 11, 9)),
    ("mu_lorentz", 1.0, 0.0, (), 1.3, 0.0, (LOR_MU,), (2, 7, 6, 7, 11, 10)),
    ("eps_lorentz", 1.5, 0.0, (LOR_EPS,), 1.0, 0.0, (), (8, 1, 1, 12, 5, 4)),
]
REAL = ("mu_debye", "mu_drude")
ALL = tuple(b[0] for b in BOXES)


def _axes():
    return tuple(np.cumsum(np.concatenate([[0.0], 1e-3 * (1.0 + 0.05 * (np.arange(n) % 3))]))
                 for n in N)


def _boxes(axes, names):
    x, y, z = axes
    out = []
    for b in BOXES:
        if b[0] in names:
            i0, j0, k0, i1, j1, k1 = b[7]
            out.append((b, (x[i0], y[j0], z[k0], x[i1], y[j1], z[k1])))
    return out


def _e0(seed, dtype, pec):
    nx, ny, nz = N
    ne = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz
    e = np.random.default_rng(seed).standard_normal(ne)
    e[pec] = 0.0
    return e.astype(dtype)


def _walls(cpml):
    return [f for f in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax") if not (cpml and f == "zmax")]


def _oracle(axes, names, dtype, steps, seed=5, cpml=False):
    x, y, z = axes
    bx = _boxes(axes, names)
    mid = om.fill_aabb(x, y, z, [(m + 1, box) for m, (_, box) in enumerate(bx)])
    col = lambda i: [1.0 if i in (1, 4) else 0.0] + [b[i] for b, _ in bx]  # noqa: E731
    me = om.mass_e(x, y, z, mid, [col(1)] * 3, om.EPS0)
    ms = om.mass_e(x, y, z, mid, [col(2)] * 3, 1.0)
    mmu = om.mass_h(x, y, z, mid, [col(4)] * 3, om.MU0)
    msm = om.mass_h(x, y, z, mid, [col(5)] * 3, 1.0)
    dt = om.courant_dt(x, y, z, 0.95)
    pec = om.bbox_walls(np.zeros(me.size, bool), *N, _walls(cpml))
    eps_poles = {m + 1: list(b[3]) for m, (b, _) in enumerate(bx) if b[3]}
    mu_poles = {m + 1: list(b[6]) for m, (b, _) in enumerate(bx) if b[6]}
    de = od.Dispersion(x, y, z, mid, eps_poles, dt, pec_e=pec) if eps_poles else None
    dh = od.Dispersion(x, y, z, mid, mu_poles, dt, side="H", excluded=mmu <= 0) if mu_poles else None
    ae, be = om.alpha_beta(me, ms, dt, pec, w=de.w_half if de else None)
    ah, bh = om.alpha_beta(mmu, msm, dt, w=dh.w_half if dh else None)
    if de is not None:
        de.bind(be, dtype)
    if dh is not None:
        dh.bind(bh, dtype)
    faces = ([oc.CPMLFace("zmax", x, y, z, dt, dtype, pec_e_flat=pec, pec_cell=np.zeros(N, bool),
                          n_layers=4)] if cpml else [])
    orc = ofd.Oracle(N, ae, be, ah, bh, dt, cpml=faces, dtype=dtype, dispersion=de, dispersion_mu=dh)
    orc.e[:] = _e0(seed, dtype, pec)
    orc.run(steps)
    return orc, dt, pec


def _solver(axes, names, precision, steps, dt, nsets=1, use_graphs=False, cpml=False):
    bx = _boxes(axes, names)
    grid = mm.GridLines(*axes)
    regions = [(mm.Material(name=b[0], epsilon=(b[1],) * 3, sigma=(b[2],) * 3, poles=b[3],
                            mu=(b[4],) * 3, sigma_m=(b[5],) * 3, poles_mu=b[6]), box)
               for b, box in bx]
    walls = _walls(cpml)
    mesh = mm.Mesh.from_grid(grid, regions=regions, pec_faces=walls)
    bcs = {f: PECBoundary(f) for f in walls}
    if cpml:
        bcs["zmax"] = CPMLBoundary("zmax", grid, thickness_cells=4)
    return FITTimeDomainSolver(mesh=mesh, boundary_conditions=bcs, nsets=nsets, total_time_steps=steps,
                               dt=dt, precision=precision, use_graphs=use_graphs)


def _seed(sol, s, seed, dtype, pec):
    sol.e[s].copy_(torch.from_numpy(layout.to_device(_e0(seed, dtype, pec), *N, "E")))


def _compare(e, h, orc, exact, tol):
    if exact:
        assert np.array_equal(e, orc.e)
        assert np.array_equal(h, orc.h)
    else:
        for got, want in ((e, orc.e), (h, orc.h)):
            err = np.linalg.norm(got.astype(np.float64) - want) / np.linalg.norm(want)
            assert err < tol, err


@pytest.mark.parametrize("precision,dtype", [("double", np.float64), ("single", np.float32)])
@pytest.mark.parametrize("names", [REAL, ALL], ids=["real_poles", "with_pairs"])
def test_mu_dispersion_matches_oracle(names, precision, dtype, gpu):
    steps = 300
    print(f"mu(w) {names} {precision}", flush=True)
    axes = _axes()
    orc, dt, pec = _oracle(axes, names, dtype, steps)
    sol = _solver(axes, names, precision, steps, dt)
    sol.setup()
    assert sol._dispersion_mu is not None
    _seed(sol, 0, 5, dtype, pec)
    sol.run()
    e, h = sol.fields(0)
    assert np.abs(orc.h).max() > 0
    exact = names == REAL or _numpy_complex_fma()
    _compare(e, h, orc, exact, 1e-12 if dtype == np.float64 else 1e-5)


def test_mu_dispersion_with_cpml_fp32(gpu):
    steps = 250
    print("mu(w) + eps(w) + CPML zmax fp32", flush=True)
    axes = _axes()
    orc, dt, pec = _oracle(axes, ALL, np.float32, steps, cpml=True)
    sol = _solver(axes, ALL, "single", steps, dt, cpml=True)
    sol.setup()
    _seed(sol, 0, 5, np.float32, pec)
    sol.run()
    e, h = sol.fields(0)
    _compare(e, h, orc, _numpy_complex_fma(), 1e-5)


def test_mu_dispersion_sets_graphs_and_resume(gpu):
    steps = 200
    print("mu(w) K=2 graphs + resume fp64", flush=True)
    axes = _axes()
    _, dt, pec = _oracle(axes, ALL, np.float64, 0)
    one = []
    for seed in (5, 9):
        sol = _solver(axes, ALL, "double", steps, dt)
        sol.setup()
        _seed(sol, 0, seed, np.float64, pec)
        sol.run()
        one.append(sol.fields(0))
    two = _solver(axes, ALL, "double", steps, dt, nsets=2, use_graphs=True)
    two.setup()
    _seed(two, 0, 5, np.float64, pec)
    _seed(two, 1, 9, np.float64, pec)
    two.run(80)
    sd = two.state_dict()
    assert "dispersion_mu" in sd["sets"][0] if "sets" in sd else "dispersion_mu" in sd
    back = _solver(axes, ALL, "double", steps, dt, nsets=2)
    back.load_state_dict(sd)
    back.run()
    two.run()
    for s in range(2):
        for sol in (two, back):
            e, h = sol.fields(s)
            assert np.array_equal(e, one[s][0])
            assert np.array_equal(h, one[s][1])
