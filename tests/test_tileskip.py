"""Dead-unit plan (CPU, torch on the host) against a brute-force census
written from the reference's rules (_tile_skip.py:136-256): dead E edges
(alpha = beta = 0), no-op H faces (alpha = 1, beta = 0) and curl-dead H faces
(four dead bounding E edges) with alpha = 1, mapped to the launch units of
the curl kernels (256 consecutive x-fastest nodes x kchunk z planes)."""

import numpy as np
import pytest
import torch

from magnelio_amd import layout
from magnelio_amd import mesh as mm
from magnelio_amd.tileskip import build_tile_plan

DIMS = (30, 20, 9)


def _coeffs(seed):
    nx, ny, nz = DIMS
    g = mm.GridLines(*(np.linspace(0, 1e-3 * n, n + 1) for n in DIMS))
    rng = np.random.default_rng(seed)
    # a thick metal wall across the guide plus seeded metal bricks
    regions = [(mm.Material.pec(), (0.0, 0.0, 0.3 * g.z[-1], g.x[-1], g.y[-1], 0.75 * g.z[-1]))]
    for _ in range(4):
        lo = rng.uniform(0, 0.6, 3) * np.array([g.x[-1], g.y[-1], g.z[-1]])
        hi = lo + rng.uniform(0.2, 0.5, 3) * np.array([g.x[-1], g.y[-1], g.z[-1]])
        regions.append((mm.Material.pec(), (*lo, *hi)))
    # magnetic loss against the xmin wall: lossy curl-dead faces stay live
    regions.append((mm.Material(name="lossy_mu", sigma_m=(80.0,) * 3),
                    (0.0, 0.0, 0.0, 0.2 * g.x[-1], g.y[-1], 0.25 * g.z[-1])))
    mesh = mm.Mesh.from_grid(g, regions=regions)
    me, mu = mm.build_m_eps(mesh), mm.build_m_mu(mesh)
    pec = mm.pec_flat(mesh.pec_mask_edges, DIMS)
    co = mm.update_coefficients(me, mu, mm.build_m_sigma(mesh), mm.build_m_sigma_m(mesh), 1e-12, pec)
    ah, bh = co.alpha_h.copy(), co.beta_h.copy()
    noop = rng.random(ah.size) < 0.05  # frozen faces (M_mu = 0 in the reference)
    ah[noop], bh[noop] = 1.0, 0.0
    return co.alpha_e, co.beta_e, ah, bh


def _census(ae, be, ah, bh, kchunk, force=()):
    nx, ny, nz = DIMS
    sh_e = [(nx, ny + 1, nz + 1), (nx + 1, ny, nz + 1), (nx + 1, ny + 1, nz)]
    sh_h = [(nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)]

    def split(v, shapes):
        out, o = [], 0
        for s in shapes:
            n = int(np.prod(s))
            out.append(v[o:o + n].reshape(s))
            o += n
        return out

    dead = (ae == 0) & (be == 0)
    live_e = ~dead
    live_e[list(force)] = True
    ex, ey, ez = split(dead, sh_e)
    cd = [ey[:, :, :-1] & ey[:, :, 1:] & ez[:, :-1, :] & ez[:, 1:, :],
          ez[:-1, :, :] & ez[1:, :, :] & ex[:, :, :-1] & ex[:, :, 1:],
          ex[:, :-1, :] & ex[:, 1:, :] & ey[:-1, :, :] & ey[1:, :, :]]
    noop = split((ah == 1) & (bh == 0), sh_h)
    unit = split(ah == 1, sh_h)
    ntp = ((nx + 1) * (ny + 1) + 255) // 256
    nkc = (nz + 1 + kchunk - 1) // kchunk
    ue, uh = set(), set()
    for comp, arr in enumerate(split(live_e, sh_e)):
        for i, j, k in zip(*np.nonzero(arr)):
            ue.add((k // kchunk) * ntp + (j * (nx + 1) + i) // 256)
    for comp in range(3):
        arr = ~(noop[comp] | (cd[comp] & unit[comp]))
        for i, j, k in zip(*np.nonzero(arr)):
            uh.add((k // kchunk) * ntp + (j * (nx + 1) + i) // 256)
    return sorted(ue), sorted(uh), dead, ntp * nkc


@pytest.mark.parametrize("kchunk", [1, 2, 3])
def test_plan_matches_census(kchunk):
    ae, be, ah, bh = _coeffs(4)
    nx, ny, nz = DIMS
    dev = [torch.from_numpy(layout.to_device(a, *DIMS, s)) for a, s in
           ((ae, "E"), (be, "E"), (ah, "H"), (bh, "H"))]
    force_ref = np.nonzero((ae == 0) & (be == 0))[0][::97][:3]  # a few element edges in metal
    plan = build_tile_plan(*dev, DIMS, kchunk, force_live_e=layout.index_to_device(force_ref, *DIMS, "E"))
    ue, uh, dead, units = _census(ae, be, ah, bh, kchunk, force_ref)
    assert plan.units == units
    assert plan.live_e.tolist() == ue and plan.live_h.tolist() == uh
    assert plan.live_e.dtype == torch.int32
    assert torch.equal(plan.zero_e, torch.from_numpy(layout.to_device(dead, *DIMS, "E")))
    # the case skips something on both sides and keeps something
    assert 0 < len(ue) < units and 0 < len(uh) < units
