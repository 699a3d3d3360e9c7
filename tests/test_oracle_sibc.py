"""Pins for the oracle's surface-impedance walls (CPU).

* the operator against a per-step numerical solve of the implicit system
  written from the impedance itself: Faraday on the face,
  M (h+ - h-) = dt R - dt sigma* h_mid - dt G [c0 h_mid + sum_p (c_p h_mid -
  (u_p+ + u_p-)/2)], each branch trapezoidal du/dt = -b u + c b h_mid
  (its voltage c s/(s + b) h = c h - u) — one face with two branches, and a
  seam face booked by two walls with their own ladders (the reference's
  gate, test_sibc_operator.py:300-387);
* Z = 0 builds no operator (the PEC path is left untouched);
* against the reference's own enumeration, fits and run of a graded cavity
  (tests/golden/sibc_cavity.npz): every wall surface (tags, components,
  indices, booked weights, dual lengths) equal, the Foster fits equal to
  1e-9, alpha_H / beta_H with the folded W bit for bit, the float64 march
  within 1e-12 and the float32 march bit for bit after 1, 100 and 300
  steps, the branch states of the step-100 checkpoint within 1e-12.
"""

import numpy as np
import pytest

import cases
from oracle import mesh as om
from oracle import sibc as osb

DT = 1e-12
D = 1e-3


def _one_face(g, tag, comp=2, flat_idx=0):
    return osb.Surface(tag, [comp], [flat_idx], [g * D * D], [1.0 / D])


def _implicit(m, dt, walls, drive, sigma_m=0.0):
    """Face state h and every wall's branch states u, advanced by solving
    the coupled linear system of each step directly."""
    br = [(g, fit.c0, b, c) for g, fit in walls for b, c in fit.branches]
    c0g = sum(g * fit.c0 for g, fit in walls)
    nb = len(br)
    h, u = 0.0, np.zeros(nb)
    out = []
    for r in drive:
        # unknowns [h+, u_1+ .. u_nb+]
        a = np.zeros((nb + 1, nb + 1))
        rhs = np.zeros(nb + 1)
        # M (h+ - h) = dt r - dt sigma (h+ + h)/2 - dt [c0g (h+ + h)/2
        #              + sum g (c (h+ + h)/2 - (u+ + u)/2)]
        a[0, 0] = m + 0.5 * dt * sigma_m + 0.5 * dt * c0g + sum(0.5 * dt * g * c for g, _, _, c in br)
        rhs[0] = m * h + dt * r - 0.5 * dt * sigma_m * h - 0.5 * dt * c0g * h
        for i, (g, _, b, c) in enumerate(br):
            rhs[0] -= 0.5 * dt * g * c * h
            a[0, 1 + i] = -0.5 * dt * g
            rhs[0] += 0.5 * dt * g * u[i]
            # (u+ - u)/dt = -b (u+ + u)/2 + c b (h+ + h)/2
            a[1 + i, 1 + i] = 1.0 / dt + 0.5 * b
            a[1 + i, 0] = -0.5 * c * b
            rhs[1 + i] = u[i] / dt - 0.5 * b * u[i] + 0.5 * c * b * h
        x = np.linalg.solve(a, rhs)
        h, u = x[0], x[1:]
        out.append(h)
    return np.array(out)


def _operator(walls_surfs, fits, m, dt, drive, sigma_m=0.0):
    dims = (2, 2, 2)
    nx, ny, nz = dims
    n_h = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1)
    op = osb.SIBC.build(walls_surfs, fits, dims, dt)
    face = int(walls_surfs[0].state_indices(dims)[0])
    mm = np.full(n_h, m)
    sm = np.full(n_h, sigma_m) + op.W
    ah, bh = om.alpha_beta(mm, sm, dt)
    op.bind(bh, np.float64)
    h = np.zeros(n_h)
    out = []
    for r in drive:
        op.save_field(h)
        rv = np.zeros(n_h)
        rv[face] = r
        h = ah * h + bh * rv
        op.update_field(h)
        out.append(h[face])
    return np.array(out)


def test_one_face_two_branches_solves_the_implicit_system():
    fit = osb.Fit(0.3, ((2e10, 0.05), (2e11, 0.12)))
    m = 1.26e-9
    drive = 1e-9 * np.random.default_rng(5).standard_normal(80)
    got = _operator([_one_face(1.0, 1)], {1: fit}, m, DT, drive)
    want = _implicit(m, DT, [(1.0, fit)], drive)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-24)
    assert np.abs(got).max() > 0


def test_seam_face_booked_by_two_walls():
    f1 = osb.Fit(0.2, ((3e10, 0.04),))
    f2 = osb.Fit(0.9, ((1e10, 0.30), (5e11, 0.08)))
    m = 1.26e-9
    drive = 1e-9 * np.random.default_rng(9).standard_normal(80)
    got = _operator([_one_face(0.5, 1), _one_face(0.5, 2)], {1: f1, 2: f2}, m, DT, drive, sigma_m=2e-4)
    want = _implicit(m, DT, [(0.5, f1), (0.5, f2)], drive, sigma_m=2e-4)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-24)


def test_a_wrong_branch_coefficient_is_caught():
    """The implicit pin is sharp: a 1 % error in one branch's c moves the
    march far beyond the gate."""
    fit = osb.Fit(0.3, ((2e10, 0.05), (2e11, 0.12)))
    bad = osb.Fit(0.3, ((2e10, 0.0505), (2e11, 0.12)))
    m = 1.26e-9
    drive = 1e-9 * np.random.default_rng(5).standard_normal(80)
    got = _operator([_one_face(1.0, 1)], {1: bad}, m, DT, drive)
    want = _implicit(m, DT, [(1.0, fit)], drive)
    assert np.max(np.abs(got - want)) > 1e-8 * np.max(np.abs(want))


def test_zero_impedance_builds_nothing():
    assert osb.SIBC.build([_one_face(1.0, 1)], {1: osb.Fit(0.0)}, (2, 2, 2), DT) is None


def test_surfaces_match_the_reference():
    f, ref, _ = cases.golden_sibc()
    got = osb.enumerate_surfaces(f["x"], f["y"], f["z"], f["material_id"],
                                 np.isin(np.arange(f["material_id"].max() + 1), f["pec_cell_ids"]),
                                 f["pec_edges"], osb.FACES)
    assert [s.tag for s in got] == [s.tag for s in ref]
    for a, b in zip(got, ref):
        assert np.array_equal(a.comp, b.comp) and np.array_equal(a.flat_idx, b.flat_idx), a.tag
        np.testing.assert_allclose(a.weight, b.weight, rtol=1e-15, atol=0)
        np.testing.assert_allclose(a.inv_l_dual, b.inv_l_dual, rtol=1e-15, atol=0)


@pytest.mark.parametrize("sigma", [5.8e5, 2.0e6])
def test_fits_match_the_reference(sigma):
    _, ref, fits = cases.golden_sibc()
    want = fits[1] if sigma == 5.8e5 else fits["zmin"]
    got = osb.fit_surface_impedance(sigma, f_lo=1e9, f_hi=1e11)
    assert abs(got.c0 - want.c0) <= 1e-9 * abs(want.c0)
    assert len(got.branches) == len(want.branches)
    np.testing.assert_allclose(np.array(got.branches), np.array(want.branches), rtol=1e-9)
    f = np.logspace(9, 11, 201)
    rs = osb.smooth_surface_impedance(f, sigma).real
    assert np.max(np.abs(got.impedance(f).real - rs) / rs) <= 1e-3


def test_coefficients_match_the_reference():
    f, _, _ = cases.golden_sibc()
    orc, sib, _ = cases.golden_sibc_oracle(np.float64)
    assert np.array_equal(orc.ah, f["double_alpha_h"])
    assert np.array_equal(orc.bh, f["double_beta_h"])


@pytest.mark.parametrize("prec,dtype", [("double", np.float64), ("single", np.float32)])
def test_march_matches_the_reference(prec, dtype):
    f, _, _ = cases.golden_sibc()
    orc, sib, _ = cases.golden_sibc_oracle(dtype)
    done = 0
    for n in (1, 100, 300):
        orc.run(n - done)
        done = n
        for comp, got in (("e", orc.e), ("h", orc.h)):
            want = f[f"{prec}_{comp}{n}"]
            if dtype == np.float32:
                assert np.array_equal(got.astype(np.float64), want), (n, comp)
            else:
                assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want), (n, comp)
        if n == 100:
            for b in sib.blocks:
                for i, u in enumerate(b.u):
                    ref = f[f"{prec}_state/sibc/{b.tag}/u{i}"]
                    scale = max(np.abs(ref).max(), 1e-300)
                    assert np.max(np.abs(u.astype(np.float64) - ref)) <= 1e-12 * scale, (b.tag, i)
