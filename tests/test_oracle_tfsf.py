"""Pins for the oracle's TF/SF plane wave (CPU).

The reference's own gate (test_plane_wave.py:80-172): a Gaussian plane
wave in an empty 24^3 PEC box, TF box inset 6 cells, run until the pulse
peak reaches the box centre; the polarisation component (grid voltage over
the edge length) reaches 1 inside the TF region within 1e-3 and stays
below 1e-4 in the SF shell, for all six propagation axes.  Any sign,
index, metric or delay error on one of the twelve faces breaks a bound.
"""

import math

import numpy as np
import pytest

from oracle import fdtd as ofd
from oracle import mesh as om
from oracle import tfsf as ot
from oracle import yee

CASES = [
    ((0, 0, 1), (1, 0, 0)),
    ((0, 0, -1), (0, 1, 0)),
    ((1, 0, 0), (0, 1, 0)),
    ((-1, 0, 0), (0, 0, 1)),
    ((0, 1, 0), (0, 0, 1)),
    ((0, -1, 0), (1, 0, 0)),
]


def tfsf_case(direction, polarization, n=24, inset=6, dtype=np.float64):
    """Oracle set-up of the reference's amplitude/leakage case."""
    L = 24e-3
    axes = [np.linspace(0.0, L, n + 1) if direction[a] >= 0 else np.linspace(-L, 0.0, n + 1)
            for a in range(3)]
    x, y, z = axes
    mid = np.zeros((n, n, n), np.int32)
    me = om.mass_e(x, y, z, mid, [[1.0]] * 3, om.EPS0)
    mm = om.mass_h(x, y, z, mid, [[1.0]] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    pec = om.bbox_walls(np.zeros(me.size, bool), n, n, n, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mm, np.zeros_like(mm), dt)
    f_max = 6e9
    corners = (tuple(a[inset] for a in axes), tuple(a[n - inset] for a in axes))
    box = ot.snap_box(x, y, z, corners)
    src = ot.PlaneWave(x, y, z, be, bh, dt, direction=direction, polarization=polarization, box=box,
                       f_max=f_max, dtype=dtype)
    centre = np.array([0.5 * (a[0] + a[-1]) for a in axes])
    n_steps = int(math.ceil((4.0 / f_max + float(np.dot(direction, centre)) / om.C0) / dt))
    orc = ofd.Oracle((n, n, n), ae, be, ah, bh, dt, sources=[src], dtype=dtype)
    return orc, n_steps, axes


@pytest.mark.parametrize("direction,polarization", CASES)
def test_tfsf_amplitude_inside_and_quiet_outside(direction, polarization):
    n, inset = 24, 6
    orc, n_steps, axes = tfsf_case(direction, polarization, n, inset)
    orc.run(n_steps)
    comp = int(np.argmax(np.abs(polarization)))
    arr = np.abs(yee.split(orc.e, n, n, n, "E")[comp]) / (24e-3 / n)
    core = slice(inset + 1, n - inset)
    inside = float(arr[core, core, core].max())
    shell = arr.copy()
    shell[inset:-inset, inset:-inset, inset:-inset] = 0.0
    outside = float(shell[1:-1, 1:-1, 1:-1].max())
    assert inside == pytest.approx(1.0, abs=1e-3), inside
    assert outside < 1e-4, outside


def test_snap_box_defaults_and_order():
    x = y = z = np.linspace(0, 1, 11)
    assert ot.snap_box(x, y, z, None) == (2, 9, 2, 9, 2, 9)
    b = ot.snap_box(x, y, z, ((0.8, None, 0.3), (0.2, 0.55, math.inf)))
    assert b == (2, 8, 2, 6, 3, 9)
