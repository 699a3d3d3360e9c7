"""GPU parity of the ADE pole-residue dielectrics (E side).

A PEC box on a graded grid holding three dispersive blocks and a lossy
dielectric, started from a seeded random H field, against the oracle's
restatement of DispersionOperator (test_solver_dispersion.py:109-158 is the
reference's own gate for the same operator):

* Debye + Drude (real poles only), FP64 and FP32: fields bit-identical;
* a Lorentz conjugate pair (complex states) added, FP64 and FP32: fields
  bit-identical when the host's NumPy rounds complex products with FMA (the
  rounding the kernel follows), else within 1e-12 / 1e-5 relative L2;
* K = 2 field sets started from different fields: each set equals its own
  single-set run bit for bit;
* graph replay equals eager launches; a state_dict round trip mid-run
  resumes to the same fields.
"""

import math

import numpy as np
import pytest
import torch

from magnelio_amd import layout
from magnelio_amd import mesh as mm
from magnelio_amd.solver import FITTimeDomainSolver
from oracle import dispersion as od
from oracle import fdtd as ofd
from oracle import mesh as om

pytestmark = pytest.mark.gpu

N = (14, 12, 10)
TAU = 1.0 / (2 * math.pi * 6e9)
WP, GAMMA = 2 * math.pi * 12e9, 2 * math.pi * 3e9
LORENTZ = od.lorentz_pair(1.8, 2 * math.pi * 15e9, 2 * math.pi * 2e9)
BOXES = [  # (name, eps_inf, sigma, poles, box in cells)
    ("debye", 2.2, 0.02, ((-1.0 / TAU, 1.5 / TAU),), (2, 2, 2, 7, 6, 5)),
    ("drude", 1.0, 0.0, ((0.0, WP * WP / GAMMA), (-GAMMA, -WP * WP / GAMMA)), (7, 6, 5, 12, 11, 9)),
    ("lorentz", 1.5, 0.0, (LORENTZ,), (2, 7, 5, 7, 11, 9)),
    ("lossy", 3.0, 0.05, (), (8, 1, 1, 12, 5, 4)),
]


def _axes():
    return tuple(np.cumsum(np.concatenate([[0.0], 1e-3 * (1.0 + 0.04 * np.arange(n) % 3)]))
                 for n in N)


def _boxes(axes, names):
    x, y, z = axes
    out = []
    for b in BOXES:
        if b[0] in names:
            i0, j0, k0, i1, j1, k1 = b[4]
            out.append((b, (x[i0], y[j0], z[k0], x[i1], y[j1], z[k1])))
    return out


def _h0(seed, dtype):
    return np.random.default_rng(seed).standard_normal(om_nh()).astype(dtype)


def om_nh():
    nx, ny, nz = N
    return (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)


def _oracle(axes, names, dtype, steps, seed=3):
    x, y, z = axes
    bx = _boxes(axes, names)
    mid = om.fill_aabb(x, y, z, [(m + 1, box) for m, (_, box) in enumerate(bx)])
    eps = [1.0] + [b[1] for b, _ in bx]
    sig = [0.0] + [b[2] for b, _ in bx]
    me = om.mass_e(x, y, z, mid, [eps] * 3, om.EPS0)
    ms = om.mass_e(x, y, z, mid, [sig] * 3, 1.0)
    mmu = om.mass_h(x, y, z, mid, [[1.0] * len(eps)] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    pec = om.bbox_walls(np.zeros(me.size, bool), *N, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    disp = od.Dispersion(x, y, z, mid, {m + 1: list(b[3]) for m, (b, _) in enumerate(bx) if b[3]},
                         dt, pec_e=pec)
    ae, be = om.alpha_beta(me, ms, dt, pec, w=disp.w_half)
    ah, bh = om.alpha_beta(mmu, np.zeros_like(mmu), dt)
    disp.bind(be, dtype)
    orc = ofd.Oracle(N, ae, be, ah, bh, dt, dtype=dtype, dispersion=disp)
    orc.h[:] = _h0(seed, dtype)
    orc.run(steps)
    return orc, dt


def _hip(axes, names, precision, steps, dt, seeds=(3,), use_graphs=False, split=None):
    bx = _boxes(axes, names)
    grid = mm.GridLines(*axes)
    regions = [(mm.Material(name=b[0], epsilon=(b[1],) * 3, sigma=(b[2],) * 3, poles=b[3]), box)
               for b, box in bx]
    mesh = mm.Mesh.from_grid(grid, regions=regions)
    dtype = np.float64 if precision == "double" else np.float32
    sol = FITTimeDomainSolver(mesh=mesh, nsets=len(seeds), total_time_steps=steps, dt=dt,
                              precision=precision, use_graphs=use_graphs)
    sol.setup()
    for s, seed in enumerate(seeds):
        sol.h[s].copy_(torch.from_numpy(layout.to_device(_h0(seed, dtype), *N, "H")))
    if split is None:
        sol.run()
    else:
        sol.run(split)
        sd = sol.state_dict()
        sol2 = FITTimeDomainSolver(mesh=mesh, nsets=len(seeds), total_time_steps=steps, dt=dt,
                                   precision=precision)
        sol2.load_state_dict(sd)
        sol2.run()
        sol = sol2
    return sol


def _numpy_complex_fma() -> bool:
    """True when NumPy's complex64 multiply rounds re = fma(ar, br, -(ai bi))
    (checked exactly with rationals on seeded operands)."""
    from fractions import Fraction

    rng = np.random.default_rng(5)
    a = (rng.standard_normal(256) + 1j * rng.standard_normal(256)).astype(np.complex64)
    b = np.complex64(complex(0.93, 0.21))
    r = (a * b).real
    plain_differs = False
    for v, got in zip(a, r):
        ar, ai, br, bi = (Fraction(float(t)) for t in (v.real, v.imag, b.real, b.imag))
        p = Fraction(float(np.float32(float(ai * bi))))
        if np.float32(float(ar * br - p)) != got:
            return False
        plain_differs |= np.float32(np.float32(float(ar * br)) - np.float32(float(p))) != got
    return bool(plain_differs)


REAL = ("debye", "drude", "lossy")
ALL = ("debye", "drude", "lorentz", "lossy")


@pytest.mark.parametrize("precision,dtype", [("double", np.float64), ("single", np.float32)])
@pytest.mark.parametrize("names", [REAL, ALL], ids=["real_poles", "with_pair"])
def test_ade_matches_oracle(names, precision, dtype, gpu):
    steps = 240
    axes = _axes()
    print(f"ADE {names} {precision}", flush=True)
    orc, dt = _oracle(axes, names, dtype, steps)
    sol = _hip(axes, names, precision, steps, dt)
    e, h = sol.fields(0)
    assert np.abs(orc.e).max() > 0
    disp = sol._dispersion
    assert disp is not None and len(disp.blocks) == len([n for n in names if n != "lossy"])
    exact = "lorentz" not in names or _numpy_complex_fma()
    if exact:
        assert np.array_equal(e, orc.e) and np.array_equal(h, orc.h)
    else:
        tol = 1e-12 if dtype == np.float64 else 1e-5
        assert np.linalg.norm(e - orc.e) <= tol * np.linalg.norm(orc.e)
        assert np.linalg.norm(h - orc.h) <= tol * np.linalg.norm(orc.h)


def test_ade_pole_states_match_oracle(gpu):
    steps = 120
    axes = _axes()
    orc, dt = _oracle(axes, ALL, np.float64, steps)
    sol = _hip(axes, ALL, "double", steps, dt)
    J = sol._dispersion.J[0].cpu().numpy()  # (pmax, n, 2)
    off = 0
    for b in orc.dispersion.blocks:
        n = b.idx.size
        for q, p in enumerate(b.poles):
            got = J[q, off:off + n, 0] + (1j * J[q, off:off + n, 1] if np.iscomplexobj(p.J) else 0)
            np.testing.assert_allclose(got, p.J, rtol=0, atol=1e-12 * np.abs(p.J).max())
        off += n


def test_ade_nsets_and_graphs(gpu):
    steps = 160
    axes = _axes()
    _, dt = _oracle(axes, ALL, np.float32, 1)
    one = [_hip(axes, ALL, "single", steps, dt, seeds=(s,)).fields(0) for s in (3, 11)]
    two = _hip(axes, ALL, "single", steps, dt, seeds=(3, 11), use_graphs=True)
    for s in (0, 1):
        e, h = two.fields(s)
        assert np.array_equal(e, one[s][0]) and np.array_equal(h, one[s][1])


def test_ade_checkpoint_resume(gpu):
    steps = 150
    axes = _axes()
    _, dt = _oracle(axes, ALL, np.float64, 1)
    full = _hip(axes, ALL, "double", steps, dt).fields(0)
    resumed = _hip(axes, ALL, "double", steps, dt, split=70).fields(0)
    assert np.array_equal(full[0], resumed[0]) and np.array_equal(full[1], resumed[1])
