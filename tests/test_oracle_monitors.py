"""Pins for the oracle's DFT monitor (CPU only).

* DFT of DC: sum_n dt exp(j w n dt) has the closed form
  dt (1 - e^{j w N dt}) / (1 - e^{j w dt});
* DFT of a sampled cosine at one of its own bin frequencies equals N dt / 2;
* cell-centre interpolation reproduces a linear field exactly at the
  cell centre (E: mean of the four edge values over the length; H: mean of
  the two face values over the dual widths) — uniform grid.
"""

import numpy as np
import pytest

from oracle import monitors as om
from oracle import yee


def test_dft_dc_closed_form():
    dt, N = 1e-12, 500
    f = np.array([1e9, 7.3e9, 20e9])
    got = om.dft_direct(np.ones(N), f, dt)
    w = 2 * np.pi * f
    want = dt * (1 - np.exp(1j * w * N * dt)) / (1 - np.exp(1j * w * dt))
    assert np.allclose(got, want, rtol=1e-10)


def test_dft_of_cosine_at_bin():
    dt, N = 1e-12, 1000
    f0 = 10 / (N * dt)
    v = np.cos(2 * np.pi * f0 * np.arange(N) * dt)
    got = om.dft_direct(v, [f0], dt)[0]
    assert got == pytest.approx(N * dt / 2, rel=1e-10)


def test_cell_centre_interpolation_of_linear_fields():
    nx, ny, nz = 4, 5, 6
    x, y, z = (np.linspace(0, 1e-3 * n, n + 1) for n in (nx, ny, nz))
    d = 1e-3
    e = np.zeros(yee.n_e(nx, ny, nz))
    h = np.zeros(yee.n_h(nx, ny, nz))
    Ex, Ey, Ez = yee.split(e, nx, ny, nz, "E")
    Hx, Hy, Hz = yee.split(h, nx, ny, nz, "H")
    # edge voltage = field * length, field E_x(y, z) = 2 + 3y + 5z (linear)
    yy = y[None, :, None]
    zz = z[None, None, :]
    Ex[...] = (2 + 3 * yy + 5 * zz) * d
    Hx[...] = 7.0 * d  # uniform H_x, dual width d on a uniform grid (boundary too)
    sl = (slice(1, 3), slice(0, 5), slice(2, 6))
    out = om.cell_centred(e, h, x, y, z, ["Ex", "Hx"], sl)
    yc = (0.5 * (y[:-1] + y[1:]))[sl[1]][None, :, None]
    zc = (0.5 * (z[:-1] + z[1:]))[sl[2]][None, None, :]
    assert np.allclose(out["Ex"], np.broadcast_to(2 + 3 * yc + 5 * zc, out["Ex"].shape), rtol=1e-12)
    assert np.allclose(out["Hx"], 7.0, rtol=1e-12)
