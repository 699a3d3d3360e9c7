"""Pins for the oracle's lumped elements (CPU).

* companion transients against closed forms, through the mini
  source-resistor circuit that mirrors the port solve (the reference's
  test_companion.py:26-130): series RL and RC step responses and the
  parallel-RC voltage converge at first order to the exponentials, the
  underdamped series RLC rings at w_d = sqrt(1/LC - (R/2L)^2);
* field coupling: a discrete port (SeriesRLC(R=Z0)) driving a PEC-walled
  gap records V + Z0 I = 2 sqrt(Z0) a(t) to 1e-12 (test_lumped_port.py:
  203-238) — a wrong sign, beta or post-update read-back breaks it.
"""

import math

import numpy as np
import pytest

from oracle import fdtd as ofd
from oracle import lumped as ol
from oracle import mesh as om


def _drive(elem, Rs, Vs, dt, n):
    i_arr, v_arr = np.empty(n), np.empty(n)
    for k in range(n):
        vh, req = elem.v_hist(dt), elem.r_eq(dt)
        i = (Vs - vh) / (Rs + req)
        v = req * i + vh
        elem.advance(i, v, dt)
        i_arr[k], v_arr[k] = i, v
    return (np.arange(n) + 1) * dt, i_arr, v_arr


def _err(factory, Rs, Vs, tau, exact, dt, use_v=False):
    n = int(round(6.0 * tau / dt))
    t, i, v = _drive(factory(), Rs, Vs, dt, n)
    return float(np.max(np.abs((v if use_v else i) - exact(t))))


def test_req_closed_forms():
    dt = 2e-12
    assert ol.SeriesRLC(R=50.0).r_eq(dt) == 50.0
    assert ol.SeriesRLC(L=1e-8).r_eq(dt) == pytest.approx(2e-8 / dt)
    assert ol.SeriesRLC(C=1e-12).r_eq(dt) == pytest.approx(dt / 2e-12)
    assert ol.SeriesRLC(R=50.0).v_hist(dt) == 0.0
    assert ol.ParallelRLC(R=50.0).r_eq(dt) == pytest.approx(50.0)


def test_series_rl_and_rc_steps():
    Vs, Rs, R = 1.0, 50.0, 50.0
    L, C = 1e-8, 1e-12
    tau = L / (Rs + R)
    i_inf = Vs / (Rs + R)
    ex = lambda t: i_inf * (1 - np.exp(-t / tau))  # noqa: E731
    ec = _err(lambda: ol.SeriesRLC(R=R, L=L), Rs, Vs, tau, ex, tau / 50)
    ef = _err(lambda: ol.SeriesRLC(R=R, L=L), Rs, Vs, tau, ex, tau / 200)
    assert ef < 0.4 * ec and ef < 5e-3 * i_inf
    tau = (Rs + R) * C
    ex = lambda t: i_inf * np.exp(-t / tau)  # noqa: E731
    ec = _err(lambda: ol.SeriesRLC(R=R, C=C), Rs, Vs, tau, ex, tau / 50)
    ef = _err(lambda: ol.SeriesRLC(R=R, C=C), Rs, Vs, tau, ex, tau / 200)
    assert ef < 0.4 * ec and ef < 5e-3 * i_inf


def test_parallel_rc_voltage():
    Vs, Rs, R, C = 1.0, 50.0, 100.0, 1e-12
    v_inf = Vs * R / (R + Rs)
    tau = Rs * R * C / (R + Rs)
    ex = lambda t: v_inf * (1 - np.exp(-t / tau))  # noqa: E731
    ec = _err(lambda: ol.ParallelRLC(R=R, C=C), Rs, Vs, tau, ex, tau / 50, use_v=True)
    ef = _err(lambda: ol.ParallelRLC(R=R, C=C), Rs, Vs, tau, ex, tau / 200, use_v=True)
    assert ef < 0.4 * ec and ef < 5e-3 * v_inf


def test_series_rlc_ringing():
    Vs, Rs, R, L, C = 1.0, 5.0, 5.0, 1e-8, 1e-12
    w0 = 1 / math.sqrt(L * C)
    a = (Rs + R) / (2 * L)
    wd = math.sqrt(w0 * w0 - a * a)
    dt = (2 * math.pi / wd) / 200
    n = 4000
    _, i, _ = _drive(ol.SeriesRLC(R=R, L=L, C=C), Rs, Vs, dt, n)
    spec = np.abs(np.fft.rfft(i - i.mean()))
    f_peak = np.fft.rfftfreq(n, dt)[np.argmax(spec)]
    assert f_peak == pytest.approx(wd / (2 * math.pi), rel=0.02)


def gap_case(dtype=np.float64, n=(12, 10, 14), element=None, Z0=50.0, steps=300):
    """A discrete port across the middle of a PEC box, Gaussian-driven."""
    x, y, z = (np.linspace(0.0, 1e-3 * m, m + 1) for m in n)
    dims = n
    mid = np.zeros(dims, np.int32)
    me = om.mass_e(x, y, z, mid, [[1.0]] * 3, om.EPS0)
    mm = om.mass_h(x, y, z, mid, [[1.0]] * 3, om.MU0)
    dt = om.courant_dt(x, y, z, 0.95)
    pec = om.bbox_walls(np.zeros(me.size, bool), *dims, ["xmin", "xmax", "ymin", "ymax", "zmin", "zmax"])
    ae, be = om.alpha_beta(me, np.zeros_like(me), dt, pec)
    ah, bh = om.alpha_beta(mm, np.zeros_like(mm), dt)
    start = (x[6], y[5], z[4])
    end = (x[6], y[5], z[9])
    _, _, flat = ol.snap_chain(x, y, z, start, end)
    port = ol.Lumped("p1", flat, dt / me[flat], element or ol.SeriesRLC(R=Z0), dt, Z0=Z0)
    f_max = 40e9
    a_t = lambda t: math.exp(-((t - 4.0 / f_max) / (2.0 / (math.pi * f_max))) ** 2)  # noqa: E731
    port.set_excitation(a_t)
    orc = ofd.Oracle(dims, ae, be, ah, bh, dt, lumped=[port], dtype=dtype)
    return orc, port, a_t, (x, y, z), (start, end), steps


def test_thevenin_invariant():
    orc, port, a_t, _, _, steps = gap_case()
    orc.run(steps)
    V, I = np.array(port.V), np.array(port.I)
    src = np.array([2 * math.sqrt(50.0) * a_t((n + 1) * orc.dt) for n in range(steps)])
    peak = np.abs(src).max()
    assert peak > 0
    assert np.abs(V + 50.0 * I - src).max() / peak < 1e-12
    assert np.abs(I).max() > 1e-6 * peak / 50.0
