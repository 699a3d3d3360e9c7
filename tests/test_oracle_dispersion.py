"""Pins for the oracle's ADE dispersion (CPU).

* Drude DC pole == conductor: a pole (a = 0, r = sigma/eps0) marches the same
  fields as the sigma-material to 1e-10 (the reference's gate,
  test_solver_dispersion.py:129-162) — exercises the W folding, the history
  subtraction and the k = 1 advance;
* each pole recursion, driven by a sampled sinusoid, settles on the bilinear
  transform of r s / (s - a) (s -> (2/dt)(1 - z^-1)/(1 + z^-1)) — pins k and
  c for real and complex poles;
* a conjugate pair stored once (weight 2, real part) equals the explicit sum
  of its two complex poles (weight 1 each) to round-off;
* H side (mu(w), _dispersion.py:45-76): a mu-Drude DC pole (a = 0,
  r = sigma*/mu0) marches the same fields as the sigma* magnetic conductor
  to 1e-10 (the mirror of the first pin: W folded into alpha_H / beta_H,
  the completion after the H curl), and a stored-once pair equals its two
  explicit poles there too.
"""

import math

import numpy as np
import pytest

from oracle import dispersion as od
from oracle import fdtd as ofd
from oracle import mesh as om

DIMS = (10, 8, 6)


def _axes():
    return tuple(np.linspace(0.0, 1e-3 * n, n + 1) for n in DIMS)


def _march(eps, sigma, poles=None, steps=200, pair_weight=None, extra_poles=None):
    x, y, z = _axes()
    mid = np.ones(DIMS, np.int32)  # one material (id 1) everywhere
    me = om.mass_e(x, y, z, mid, [[1.0, eps]] * 3, om.EPS0)
    ms = om.mass_e(x, y, z, mid, [[0.0, sigma]] * 3, 1.0)
    mm = om.mass_h(x, y, z, mid, [[1.0, 1.0]] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    pec = om.bbox_walls(np.zeros(me.size, bool), *DIMS, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    disp = None
    w = None
    if poles is not None:
        disp = od.Dispersion(x, y, z, mid, {1: poles}, dt, pec_e=pec, pair_weight=pair_weight)
        w = disp.w_half
    ae, be = om.alpha_beta(me, ms, dt, pec, w=w)
    ah, bh = om.alpha_beta(mm, np.zeros_like(mm), dt)
    if disp is not None:
        disp.bind(be, np.float64)
    orc = ofd.Oracle(DIMS, ae, be, ah, bh, dt, dispersion=disp)
    orc.h[:] = np.random.default_rng(7).standard_normal(orc.h.size)
    orc.run(steps)
    return orc.e.copy(), orc.h.copy()


def test_drude_dc_pole_equals_conductor():
    sigma = 0.05
    e_s, h_s = _march(2.0, sigma)
    e_a, h_a = _march(2.0, 0.0, poles=[(0.0, sigma / om.EPS0)])
    assert np.abs(e_s).max() > 0
    assert np.abs(e_a - e_s).max() < 1e-10 * np.abs(e_s).max()
    assert np.abs(h_a - h_s).max() < 1e-10 * np.abs(h_s).max()


@pytest.mark.parametrize("a,r", [(-2 * math.pi * 3e9, 4e10),
                                 (complex(-2e9, 2 * math.pi * 8e9), complex(1e9, -6e10))])
def test_recursion_is_bilinear_transform(a, r):
    dt = 1e-12
    f = 5e9
    w = 2 * math.pi * f
    n = 40000
    p = od.Pole(a, r, dt, 1)
    e = np.cos(w * np.arange(n + 1) * dt)
    J = np.empty(n + 1, complex)
    J[0] = 0.0
    for m in range(n):
        p.J *= p.k
        p.J += p.c * (e[m + 1] - e[m])
        J[m + 1] = p.J[0]
    t = np.arange(n + 1)[-4000:] * dt
    A = np.stack([np.exp(1j * w * t), np.exp(-1j * w * t)], axis=1)
    coef, *_ = np.linalg.lstsq(A, J[-4000:], rcond=None)

    def H(sign):
        zinv = np.exp(-1j * sign * w * dt)
        s = (2 / dt) * (1 - zinv) / (1 + zinv)
        return r * s / (s - a)

    assert abs(coef[0] - H(+1) / 2) < 1e-9 * abs(H(+1))
    assert abs(coef[1] - H(-1) / 2) < 1e-9 * abs(H(-1))


def test_conjugate_pair_stored_once_equals_two_poles():
    a, r = od.lorentz_pair(1.5, 2 * math.pi * 20e9, 2 * math.pi * 2e9)
    e1, h1 = _march(2.0, 0.0, poles=[(a, r)])
    e2, h2 = _march(2.0, 0.0, poles=[(a, r), (a.conjugate(), r.conjugate())], pair_weight=1.0)
    assert np.abs(e1).max() > 0
    assert np.abs(e1 - e2).max() < 1e-12 * np.abs(e1).max()
    assert np.abs(h1 - h2).max() < 1e-12 * np.abs(h1).max()


def test_w_positive_for_passive_pole():
    x, y, z = _axes()
    mid = np.ones(DIMS, np.int32)
    a, r = od.lorentz_pair(1.5, 2 * math.pi * 20e9, 2 * math.pi * 2e9)
    disp = od.Dispersion(x, y, z, mid, {1: [(a, r), (-1e10, 3e10)]}, 1e-12)
    assert (disp.W[disp.blocks[0].idx] > 0).all()


def _march_mu(mu, sigma_m, poles=None, steps=200, pair_weight=None):
    x, y, z = _axes()
    mid = np.ones(DIMS, np.int32)
    me = om.mass_e(x, y, z, mid, [[1.0, 2.0]] * 3, om.EPS0)
    mmu = om.mass_h(x, y, z, mid, [[1.0, mu]] * 3, om.MU0)
    msm = om.mass_h(x, y, z, mid, [[0.0, sigma_m]] * 3, 1.0)
    dt = om.courant_dt(x, y, z, 0.95) / np.sqrt(mu)
    pec = om.bbox_walls(np.zeros(me.size, bool), *DIMS, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    disp = None
    w = None
    if poles is not None:
        disp = od.Dispersion(x, y, z, mid, {1: poles}, dt, pair_weight=pair_weight, side="H",
                             excluded=mmu <= 0)
        w = disp.w_half
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mmu, msm, dt, w=w)
    if disp is not None:
        disp.bind(bh, np.float64)
    orc = ofd.Oracle(DIMS, ae, be, ah, bh, dt, dispersion_mu=disp)
    orc.e[:] = np.random.default_rng(8).standard_normal(orc.e.size) * (~pec)
    orc.run(steps)
    return orc.e.copy(), orc.h.copy()


def test_mu_drude_dc_pole_equals_magnetic_conductor():
    sigma_m = 4e3  # ohm/m
    e_s, h_s = _march_mu(1.5, sigma_m)
    e_a, h_a = _march_mu(1.5, 0.0, poles=[(0.0, sigma_m / om.MU0)])
    assert np.abs(h_s).max() > 0
    # the magnetic conductor damps: the run is not trivially lossless
    e_0, _ = _march_mu(1.5, 0.0)
    assert np.abs(e_s - e_0).max() > 1e-3 * np.abs(e_0).max()
    assert np.abs(e_a - e_s).max() < 1e-10 * np.abs(e_s).max()
    assert np.abs(h_a - h_s).max() < 1e-10 * np.abs(h_s).max()


def test_mu_conjugate_pair_stored_once_equals_two_poles():
    a, r = od.lorentz_pair(0.8, 2 * math.pi * 20e9, 2 * math.pi * 2e9)
    e1, h1 = _march_mu(1.2, 0.0, poles=[(a, r)])
    e2, h2 = _march_mu(1.2, 0.0, poles=[(a, r), (a.conjugate(), r.conjugate())], pair_weight=1.0)
    assert np.abs(h1).max() > 0
    assert np.abs(e1 - e2).max() < 1e-12 * np.abs(e1).max()
    assert np.abs(h1 - h2).max() < 1e-12 * np.abs(h1).max()


def test_h_side_blocks_follow_face_owner_and_frozen_faces():
    """Faces owned by the dispersive cells (the clamped lookup of
    build_M_sigma_m) carry states, frozen faces (M_mu <= 0) none, and
    g = MU0 A_primal / L_dual equals M_mu / mu_r there."""
    x, y, z = _axes()
    mid = np.zeros(DIMS, np.int32)
    mid[2:6, 1:5, 1:4] = 1
    mmu = om.mass_h(x, y, z, mid, [[1.0, 2.5]] * 3, om.MU0)
    frozen = np.zeros(mmu.size, bool)
    frozen[::7] = True
    disp = od.Dispersion(x, y, z, mid, {1: [(-1e10, 3e10)]}, 1e-12, side="H", excluded=frozen)
    blk = disp.blocks[0]
    own = om.mass_h(x, y, z, mid, [[0.0, 1.0]] * 3, 1.0) > 0
    assert np.array_equal(np.sort(blk.idx), np.nonzero(own & ~frozen)[0])
    np.testing.assert_allclose(blk.g, mmu[blk.idx] / 2.5, rtol=1e-14)
