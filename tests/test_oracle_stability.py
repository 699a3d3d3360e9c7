"""Pins for the oracle's spectral CFL limit (CPU): the closed form of the
largest eigenvalue of the discrete Maxwell operator on a uniform PEC box,
the Gershgorin bound above it, and a graded dielectric mesh whose limit
lies between the bound and the geometric Courant step."""

import math

import numpy as np

from oracle import mesh as om
from oracle import stability as ost


def _box(dims, d, mid=None, eps=(1.0,)):
    x, y, z = (np.linspace(0.0, dd * n, n + 1) for n, dd in zip(dims, d))
    mid = np.zeros(dims, np.int32) if mid is None else mid
    me = om.mass_e(x, y, z, mid, [list(eps)] * 3, om.EPS0)
    mu = om.mass_h(x, y, z, mid, [[1.0] * len(eps)] * 3, om.MU0)
    pec = om.bbox_walls(np.zeros(me.size, bool), *dims, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    return (x, y, z), me, mu, pec


def test_uniform_box_closed_form():
    dims, d = (9, 7, 6), (1e-3, 0.8e-3, 1.2e-3)
    _, me, mu, pec = _box(dims, d)
    lam, upper = ost.lambda_max(dims, me, mu, pec)
    exact = ost.box_lambda_max(dims, d, om.EPS0, om.MU0)
    assert abs(lam - exact) < 1e-8 * exact
    assert upper >= exact * (1 - 1e-12)


def test_dielectric_box_between_bounds():
    dims, d = (10, 8, 7), (1e-3, 1e-3, 1e-3)
    mid = np.zeros(dims, np.int32)
    mid[3:7, 2:6, 2:5] = 1
    axes, me, mu, pec = _box(dims, d, mid, eps=(1.0, 6.0))
    lam, upper = ost.lambda_max(dims, me, mu, pec)
    dt = 0.95 * 2.0 / math.sqrt(lam)
    assert lam <= upper * (1 + 1e-12)
    # the spectral step is never below the geometric Courant step
    assert dt >= om.courant_dt(*axes, 0.95)
